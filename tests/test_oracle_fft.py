"""Pins the oracle's FFT and H application (the reference's FFTW calls,
filter/hamiltonian.c:27-44) against numpy.fft and closed forms."""
import numpy as np
import pytest


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 11, 12, 16, 18, 20, 24, 25, 27, 30, 32,
                               36, 48, 49, 64, 80, 96, 121, 128, 160])
def test_fft1d_vs_numpy(oracle, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    fw = oracle.fft1d(x, -1)
    bw = oracle.fft1d(x, +1)
    scale = np.abs(np.fft.fft(x)).max()
    assert np.abs(fw - np.fft.fft(x)).max() <= 1e-13 * scale
    assert np.abs(bw - np.fft.ifft(x) * n).max() <= 1e-13 * scale


@pytest.mark.parametrize("shape", [(4, 6, 8), (12, 10, 9), (16, 16, 16), (20, 24, 36)])
def test_fft3d_vs_numpy(oracle, shape):
    rng = np.random.default_rng(7)
    v = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    a = oracle.fft3d(v, -1)
    b = np.fft.fftn(v)
    assert np.abs(a - b).max() <= 1e-13 * np.abs(b).max()
    c = oracle.fft3d(v, +1)
    assert np.abs(c - np.fft.ifftn(v) * v.size).max() <= 1e-13 * np.abs(b).max()


def _plane_wave(grid, d, m):
    nx, ny, nz = grid
    x = np.arange(nx) * d[0]
    y = np.arange(ny) * d[1]
    z = np.arange(nz) * d[2]
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    k = [2 * np.pi * m[a] / (grid[a] * d[a]) for a in range(3)]
    return np.exp(1j * (k[0] * X + k[1] * Y + k[2] * Z)).reshape(-1), k


@pytest.mark.parametrize("m", [(0, 0, 0), (1, 0, 0), (0, 3, -2), (-5, 4, 7), (11, -11, 5)])
def test_plane_wave_eigenvalue(oracle, m):
    """e^{ik.r} is an eigenvector of T + V0 with eigenvalue min(k^2/2, 10) + V0
    (ksqr of filter/init.c:124-144 with its 10 Ha clamp)."""
    grid, d, V0 = (24, 20, 16), (0.5, 0.55, 0.6), -0.37
    psi, k = _plane_wave(grid, d, m)
    ksq = oracle.ksqr(*grid, *d)
    pot = np.full(psi.size, V0)
    hpsi = oracle.hamiltonian(grid, ksq, pot, psi)
    lam = min(0.5 * sum(q * q for q in k), 10.0) + V0
    assert np.abs(hpsi - lam * psi).max() < 1e-11


def test_ksqr_closed_form(oracle):
    grid, d = (10, 12, 14), (0.7, 0.8, 0.9)
    ksq = oracle.ksqr(*grid, *d, 3.0).reshape(grid[2], grid[1], grid[0])
    N = np.prod(grid)
    kk = [2 * np.pi * np.fft.fftfreq(grid[a], d[a]) for a in range(3)]
    KZ, KY, KX = np.meshgrid(kk[2], kk[1], kk[0], indexing="ij")
    ref = np.minimum(0.5 * (KX ** 2 + KY ** 2 + KZ ** 2), 3.0) / N
    assert np.allclose(ksq, ref, rtol=1e-14, atol=0)


def test_hamiltonian_vs_numpy(oracle):
    grid, d = (12, 10, 8), (0.6, 0.7, 0.8)
    rng = np.random.default_rng(3)
    N = np.prod(grid)
    pot = rng.standard_normal(N)
    psi = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    ksq = oracle.ksqr(*grid, *d)
    got = oracle.hamiltonian(grid, ksq, pot, psi)
    sh = (grid[2], grid[1], grid[0])
    t = np.fft.ifftn(np.fft.fftn(psi.reshape(sh)) * ksq.reshape(sh)) * N
    ref = t.reshape(-1) + pot * psi
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()
