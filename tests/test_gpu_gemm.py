"""Matrix-core contractions (projected H / overlaps, Ritz rotation;
filter/Hmat.c:30-46,61-78) against numpy on device buffers."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [(1, 1), (4, 3), (16, 16), (17, 33), (64, 64), (100, 37)]


@pytest.mark.parametrize("m,mb", CASES)
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_gram(sep, m, mb, prec):
    import torch
    print("case gram", m, mb, prec, flush=True)
    P = sep.SEP_PREC_FP64 if prec == "fp64" else sep.SEP_PREC_FP32
    dt = torch.float64 if prec == "fp64" else torch.float32
    with sep.Context(24, 16, 16, 1.0, 1.0, 1.0) as ctx:
        N = ctx.N
        g = torch.Generator().manual_seed(m * 100 + mb)
        A = torch.randn(m, N, generator=g, dtype=torch.float64)
        B = torch.randn(mb, N, generator=g, dtype=torch.float64)
        Ad, Bd = A.to(dt).to("cuda:0"), B.to(dt).to("cuda:0")
        C = ctx.gram_dev(Ad.data_ptr(), m, Bd.data_ptr(), mb, P)
    ref = (A.to(dt).double() @ B.to(dt).double().T).numpy()
    tol = 1e-12 if prec == "fp64" else 2e-5
    assert np.abs(C - ref).max() <= tol * np.sqrt(N) * 4, np.abs(C - ref).max()


@pytest.mark.parametrize("m,mo", CASES)
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_rotate(sep, m, mo, prec):
    import torch
    print("case rotate", m, mo, prec, flush=True)
    P = sep.SEP_PREC_FP64 if prec == "fp64" else sep.SEP_PREC_FP32
    dt = torch.float64 if prec == "fp64" else torch.float32
    with sep.Context(24, 16, 16, 1.0, 1.0, 1.0) as ctx:
        N = ctx.N
        g = torch.Generator().manual_seed(m * 7 + mo)
        A = torch.randn(m, N, generator=g, dtype=torch.float64)
        M = np.random.default_rng(m + mo).standard_normal((m, mo))
        Ad = A.to(dt).to("cuda:0")
        Od = torch.zeros(mo, N, dtype=dt, device="cuda:0")
        ctx.rotate_dev(Ad.data_ptr(), m, M, Od.data_ptr(), P)
        got = Od.double().cpu().numpy()
    ref = M.T @ A.to(dt).double().numpy()
    tol = 1e-13 if prec == "fp64" else 3e-6
    assert np.abs(got - ref).max() <= tol * np.sqrt(m) * 8, np.abs(got - ref).max()
