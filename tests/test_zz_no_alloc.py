"""sep_apply_h_dev_noalloc (the allocation-free variant of sep_apply_h_dev)
through the C ABI: the same H.psi as sep_apply_h_dev, bit for bit, on two
calls in a row with the second one smaller, in both precisions; sizes beyond
the static scratch are refused with SEP_ERR_ARG."""
import ctypes

import numpy as np
import pytest

SEP_OK, SEP_ERR_ARG = 0, 1  # include/sep.h

pytestmark = pytest.mark.gpu


def _fn(sep):
    L = sep.lib()
    f = L.sep_apply_h_dev_noalloc
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long, ctypes.c_int]
    f.restype = ctypes.c_int
    g = L.sep_apply_h_dev
    g.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long, ctypes.c_int]
    g.restype = ctypes.c_int
    return f, g


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_noalloc_matches_apply_h_dev(sep, prec):
    import torch
    f, g = _fn(sep)
    P = sep.SEP_PREC_FP64 if prec == "fp64" else sep.SEP_PREC_FP32
    dt = torch.complex128 if prec == "fp64" else torch.complex64
    grid = (48, 48, 48)
    with sep.Context(*grid, 0.6, 0.6, 0.6) as ctx:
        N = ctx.N
        rng = np.random.default_rng(3)
        ctx.set_potential(rng.uniform(-1.0, 0.5, N))
        for batch in (5, 2):  # second call smaller than the first
            print("case noalloc", prec, "batch", batch, flush=True)
            x = torch.randn(batch, N, dtype=dt, generator=torch.Generator().manual_seed(batch))
            xd = x.to("cuda:0")
            ref = torch.empty_like(xd)
            got = torch.full_like(xd, float("nan"))
            assert g(ctx._h, xd.data_ptr(), ref.data_ptr(), batch, P) == SEP_OK
            assert f(ctx._h, xd.data_ptr(), got.data_ptr(), batch, P) == SEP_OK
            torch.cuda.synchronize()
            assert torch.equal(got, ref), (got - ref).abs().max().item()
        # beyond the static scratch: refused, nothing launched
        too_many = (32 << 20) // (N * (16 if prec == "fp64" else 8)) + 1
        big = torch.zeros(too_many, N, dtype=dt, device="cuda:0")
        assert f(ctx._h, big.data_ptr(), big.data_ptr(), too_many, P) == SEP_ERR_ARG
        assert "scratch" in sep.lib().sep_last_error(ctx._h).decode()
