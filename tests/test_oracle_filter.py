"""Closed forms that pin the oracle's spectral bound (filter/energy.c:53-83)
and Newton filter (filter/filter.c:59-83, filter/hnorm.c:11-25) without the
reference binary: with a constant potential V0 every plane wave is an
eigenvector of H with eigenvalue min(k^2/2, 10) + V0."""
import numpy as np
import pytest


def _grid(oracle, n=(12, 10, 8), d=(0.7, 0.8, 0.9)):
    ksq = oracle.ksqr(*n, *d)
    return n, d, ksq


def test_energy_range_constant_potential(oracle):
    """Power iteration on H = T + V0 converges to the kinetic clamp: on a grid
    whose smallest non-zero k already has k^2/2 > 10 the spectrum is {V0,
    10 + V0}, so Emax = 10 + V0 and dE = 1.1 (Emax - Vmin) = 11 (energy.c:76)."""
    n, d, ksq = _grid(oracle, (4, 4, 4), (0.2, 0.25, 0.3))
    V0 = -0.37
    V = np.full(int(np.prod(n)), V0)
    dE, emax, it = oracle.energy_range(n, ksq, V, d[0] * d[1] * d[2], V0)
    assert abs(emax - (10.0 + V0)) < 1e-5
    assert abs(dE - 1.1 * (emax - V0)) < 1e-12
    assert 1 <= it <= 500


@pytest.mark.parametrize("m", [(0, 0, 0), (1, 2, 0), (3, -1, 2), (6, 5, 4)])
def test_filter_plane_wave_closed_form(oracle, m):
    """A real plane wave cos(k.r) passes through the Newton filter as the scalar
    polynomial p(x) = sum_j a_j prod_{l<j} (x - z_l), x = dE_1 (eps - Vmin) - 2,
    evaluated here in numpy from the coefficients alone."""
    n, d, ksq = _grid(oracle)
    nx, ny, nz = n
    V0 = 0.25
    V = np.full(nx * ny * nz, V0)
    L = [n[i] * d[i] for i in range(3)]
    k = [2 * np.pi * m[i] / L[i] for i in range(3)]
    eps = min(0.5 * (k[0] ** 2 + k[1] ** 2 + k[2] ** 2), 10.0)
    z, y, x = np.meshgrid(np.arange(nz) * d[2], np.arange(ny) * d[1], np.arange(nx) * d[0],
                          indexing="ij")
    wave = np.cos(k[0] * x + k[1] * y + k[2] * z).reshape(-1)
    nc, ms = 40, 3
    dE = 11.0
    zn, _ = oracle.leja_points(nc)
    el = np.array([V0 + 0.3, V0 + 1.0, V0 + 2.5])
    dt = (nc / (2.5 * dE)) ** 2
    an = oracle.newton_coeffs(nc, el, zn, dE, V0, dt)
    out = oracle.filter_state(n, ksq, V, an, zn, 4.0 / dE, V0, wave)
    xs = 4.0 / dE * (eps + V0 - V0) - 2.0
    prod, p = 1.0, np.zeros(ms)
    for j in range(nc):
        p += an[:, j] * prod
        prod *= xs - zn[j]
    for ie in range(ms):
        assert np.abs(out[ie] - p[ie] * wave).max() < 1e-10 * max(1.0, abs(p[ie]))
    # and the polynomial approximates the Gaussian filter (40 Newton terms)
    E = (xs + 2.0) * dE / 4.0 + V0
    gauss = np.sqrt(dt / np.pi) * np.exp(-dt * (E - el) ** 2)
    assert np.abs(p - gauss).max() < 1e-3 * np.sqrt(dt / np.pi)
