"""H.psi (filter/hamiltonian.c:11-44, hnorm.c:11-39) on the GPU against the
CPU oracle on seeded random complex vectors (FP64: <= 1e-12 relative)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GRIDS = [(16, 16, 16), (24, 24, 24), (36, 36, 36), (64, 64, 64), (32, 24, 16), (48, 48, 32),
         (96, 96, 96), (128, 128, 128), (80, 64, 64)]


def _setup(oracle, grid, seed):
    nx, ny, nz = grid
    rng = np.random.default_rng(seed)
    d = (0.7, 0.8, 0.9)
    ksq = oracle.ksqr(nx, ny, nz, *d)
    pot = rng.uniform(-1.5, 1.0, nx * ny * nz)
    return d, ksq, pot, rng


@pytest.mark.parametrize("grid", GRIDS)
def test_apply_h_fp64(sep, oracle, grid):
    print("case apply_h_fp64", grid, flush=True)
    nx, ny, nz = grid
    d, ksq, pot, rng = _setup(oracle, grid, 11)
    N = nx * ny * nz
    batch = 3 if N <= 300000 else 1
    psi = rng.standard_normal((batch, N)) + 1j * rng.standard_normal((batch, N))
    with sep.Context(nx, ny, nz, *d) as ctx:
        ctx.set_potential(pot)
        got = ctx.apply_h(psi)
    ref = oracle.hamiltonian(grid, ksq, pot, psi)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 1e-12, err


@pytest.mark.parametrize("grid", [(24, 24, 24), (64, 64, 64)])
def test_apply_h_line_kernels(sep, oracle, grid, monkeypatch):
    """The row/strided-line fallback (used when a plane does not fit LDS)."""
    print("case apply_h_lines", grid, flush=True)
    monkeypatch.setenv("SEP_FORCE_LINE_KERNELS", "1")
    nx, ny, nz = grid
    d, ksq, pot, rng = _setup(oracle, grid, 12)
    N = nx * ny * nz
    psi = rng.standard_normal((2, N)) + 1j * rng.standard_normal((2, N))
    with sep.Context(nx, ny, nz, *d) as ctx:
        ctx.set_potential(pot)
        got = ctx.apply_h(psi)
    ref = oracle.hamiltonian(grid, ksq, pot, psi)
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-12


def test_hnorm_fp64(sep, oracle):
    """Newton step psi_j = dE_1 H psi - (2 + z + Vmin dE_1) psi (hnorm.c:19-24)."""
    grid = (32, 32, 32)
    nx, ny, nz = grid
    d, ksq, pot, rng = _setup(oracle, grid, 13)
    N = nx * ny * nz
    psi = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    zm1, vmin, dE_1 = 0.37, pot.min(), 4.0 / 13.0
    with sep.Context(nx, ny, nz, *d) as ctx:
        ctx.set_potential(pot)
        got = ctx.hnorm(psi, zm1, vmin, dE_1)
    h = oracle.hamiltonian(grid, ksq, pot, psi)
    ref = dE_1 * h - (2.0 + zm1 + vmin * dE_1) * psi
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-12


@pytest.mark.parametrize("grid", [(36, 36, 36), (64, 64, 64), (128, 128, 128)])
def test_apply_h_fp32(sep, oracle, grid):
    """complex64 H.psi through the device-pointer entry (torch buffers)."""
    import torch
    print("case apply_h_fp32", grid, flush=True)
    nx, ny, nz = grid
    d, ksq, pot, rng = _setup(oracle, grid, 14)
    N = nx * ny * nz
    psi = rng.standard_normal((2, N)) + 1j * rng.standard_normal((2, N))
    with sep.Context(nx, ny, nz, *d) as ctx:
        ctx.set_potential(pot)
        x = torch.from_numpy(psi.astype(np.complex64)).to("cuda:0")
        y = torch.empty_like(x)
        ctx.apply_h_dev(x.data_ptr(), y.data_ptr(), 2, sep.SEP_PREC_FP32)
        torch.cuda.synchronize()
        got = y.cpu().numpy()
    ref = oracle.hamiltonian(grid, ksq, pot, psi)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 2e-6, err
