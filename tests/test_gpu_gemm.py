"""Matrix-core contractions (projected H / overlaps, Ritz rotation;
filter/Hmat.c:30-46,61-78, the Gram of ortho.c:32) against numpy on device
buffers: both precisions, tile-ragged shapes and the 512-state shape of the
production subspace.  Contraction lengths that are not a multiple of the
16-wide MFMA group (the element-wise tail) occur in the k x k products of
the Rayleigh-Ritz stage and are covered by tests/test_gpu_refine.py."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [(1, 1), (4, 3), (16, 16), (17, 33), (64, 64), (100, 37), (129, 200), (512, 512)]
GRIDS = [(24, 16, 16), (36, 36, 36)]


@pytest.mark.parametrize("grid", GRIDS)
@pytest.mark.parametrize("m,mb", CASES)
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_gram(sep, grid, m, mb, prec):
    import torch
    print("case gram", grid, m, mb, prec, flush=True)
    P = sep.SEP_PREC_FP64 if prec == "fp64" else sep.SEP_PREC_FP32
    dt = torch.float64 if prec == "fp64" else torch.float32
    with sep.Context(*grid, 1.0, 1.0, 1.0) as ctx:
        N = ctx.N
        g = torch.Generator().manual_seed(m * 100 + mb)
        A = torch.randn(m, N, generator=g, dtype=torch.float64)
        B = torch.randn(mb, N, generator=g, dtype=torch.float64)
        Ad, Bd = A.to(dt).to("cuda:0"), B.to(dt).to("cuda:0")
        C = ctx.gram_dev(Ad.data_ptr(), m, Bd.data_ptr(), mb, P)
    ref = (A.to(dt).double() @ B.to(dt).double().T).numpy()
    # FP64: sums of N products of unit normals, rounding ~ eps sqrt(N);
    # FP32: the MFMA accumulates each chunk in FP32
    tol = 4e-15 if prec == "fp64" else 2e-6
    err = np.abs(C - ref).max()
    print("  max err", err, flush=True)
    assert err <= tol * N, err


@pytest.mark.parametrize("grid", GRIDS)
@pytest.mark.parametrize("m,mo", CASES)
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_rotate(sep, grid, m, mo, prec):
    import torch
    print("case rotate", grid, m, mo, prec, flush=True)
    P = sep.SEP_PREC_FP64 if prec == "fp64" else sep.SEP_PREC_FP32
    dt = torch.float64 if prec == "fp64" else torch.float32
    with sep.Context(*grid, 1.0, 1.0, 1.0) as ctx:
        N = ctx.N
        g = torch.Generator().manual_seed(m * 7 + mo)
        A = torch.randn(m, N, generator=g, dtype=torch.float64)
        M = np.random.default_rng(m + mo).standard_normal((m, mo))
        Ad = A.to(dt).to("cuda:0")
        Od = torch.zeros(mo, N, dtype=dt, device="cuda:0")
        ctx.rotate_dev(Ad.data_ptr(), m, M, Od.data_ptr(), P)
        got = Od.double().cpu().numpy()
    Mr = M if prec == "fp64" else M.astype(np.float32).astype(np.float64)
    ref = Mr.T @ A.to(dt).double().numpy()
    tol = 4e-15 if prec == "fp64" else 2e-6
    err = np.abs(got - ref).max()
    print("  max err", err, flush=True)
    assert err <= tol * m, err


@pytest.mark.parametrize("mo", [17, 65, 100])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_rotate_rows_independent_of_output_count(sep, mo, prec):
    """A state's rotated row does not depend on how many states follow it:
    the waves skip the 16-row sub-tiles past the last output state without
    changing any kept state's MFMA sequence, so the first mo rows of a
    rotation to 128 states equal the rotation to mo states bit for bit."""
    import torch
    print("case rotate tail", mo, prec, flush=True)
    P = sep.SEP_PREC_FP64 if prec == "fp64" else sep.SEP_PREC_FP32
    dt = torch.float64 if prec == "fp64" else torch.float32
    m = 195
    with sep.Context(24, 16, 16, 1.0, 1.0, 1.0) as ctx:
        N = ctx.N
        g = torch.Generator().manual_seed(mo)
        Ad = torch.randn(m, N, generator=g, dtype=torch.float64).to(dt).to("cuda:0")
        Mpad = np.random.default_rng(mo).standard_normal((m, 128))
        out = []
        for M in (np.ascontiguousarray(Mpad[:, :mo]), Mpad):
            Od = torch.zeros(M.shape[1], N, dtype=dt, device="cuda:0")
            ctx.rotate_dev(Ad.data_ptr(), m, M, Od.data_ptr(), P)
            out.append(Od[:mo].cpu().numpy())
    assert np.abs(out[0]).max() > 0
    assert np.array_equal(out[0], out[1])
