"""Filter lanes (sep_set_filter_lanes): the random states split into groups
of whole complex vectors that step on streams of their own give bitwise the
filtered states of one lane -- odd and even state counts, uneven groups,
both precisions, both kernel families, three ring blocks (ring depth
bounded to 256 by the test hook SEP_RING_DEPTH).  Also the
HBM copy probe (sep_bench_copy) that the bench reports beside the peak."""
import numpy as np
import pytest

from tests.test_gpu_fused import _setup

pytestmark = pytest.mark.gpu


def _filter(sep, d, psi0, nc, ms, prec, lanes_list):
    n = d["nx"]
    ctx = sep.Context(n, n, n, d["dx"], d["dy"], d["dz"], d["xmin"], d["ymin"], d["zmin"])
    try:
        ctx.build_potential(d["xyz"], d["type"], d["npot"], d["vr"], d["pot"], d["dr"],
                            expot=d["expot"])
        dE, _, _ = ctx.energy_range()
        zn, _ = sep.leja_points(nc)
        an = sep.newton_coeffs(zn, np.linspace(d["Elmin"], d["Elmax"], ms), dE, ctx.vmin,
                               (nc / (2.5 * dE)) ** 2)
        out = []
        for lanes in lanes_list:
            ctx.set_filter_lanes(lanes)
            states, _ = ctx.filter(psi0, an, zn, ctx.vmin, dE, prec)
            out.append(states)
        return out
    finally:
        ctx.close()


@pytest.mark.parametrize("family", ["plane", "lines"])
@pytest.mark.parametrize("prec", ["fp32", "fp64"])
@pytest.mark.parametrize("n,ns", [(32, 7), (48, 8), (64, 5)])
def test_lanes_bitwise(sep, monkeypatch, family, prec, n, ns):
    monkeypatch.setenv("SEP_RING_DEPTH", "256")
    if family == "lines":
        monkeypatch.setenv("SEP_FORCE_LINE_KERNELS", "1")
    P = sep.SEP_PREC_FP32 if prec == "fp32" else sep.SEP_PREC_FP64
    nc, ms = 600, 3
    d, psi0 = _setup(sep, n, 1.2 + n / 64.0, ns, nc, ms)
    print("case lanes", family, prec, n, ns, flush=True)
    one, auto, two, three, four = _filter(sep, d, psi0, nc, ms, P, [1, 0, 2, 3, 4])
    assert np.isfinite(one).all() and np.abs(one).max() > 0.0
    for name, got in (("auto", auto), ("2", two), ("3", three), ("4", four)):
        print("lanes", name, "bitwise", np.array_equal(got, one), flush=True)
        assert np.array_equal(got, one)


def test_lanes_argument_checks(sep):
    ctx = sep.Context(16, 16, 16, 0.5, 0.5, 0.5)
    try:
        for bad in (-1, sep.SEP_MAX_LANES + 1):
            with pytest.raises(sep.SepError):
                ctx.set_filter_lanes(bad)
        ctx.set_filter_lanes(sep.SEP_MAX_LANES)
        ctx.set_filter_lanes(0)
    finally:
        ctx.close()


def test_copy_probe(sep):
    """sep_bench_copy: a plain copy's rate is positive and below the nominal
    8 TB/s peak; sizes under 1 MiB are refused."""
    ctx = sep.Context(16, 16, 16, 0.5, 0.5, 0.5)
    try:
        gbs = ctx.bench_copy(1 << 30, 2)
        print("copy probe GB/s", gbs, flush=True)
        assert 100.0 < gbs < 8000.0
        with pytest.raises(sep.SepError):
            ctx.bench_copy(1024, 1)
    finally:
        ctx.close()


def test_trim_then_filter_again(sep):
    """sep_trim frees the ring; the next filter allocates it again and gives
    bitwise the same states."""
    nc, ms = 300, 3
    d, psi0 = _setup(sep, 32, 1.7, 4, nc, ms)
    n = d["nx"]
    ctx = sep.Context(n, n, n, d["dx"], d["dy"], d["dz"], d["xmin"], d["ymin"], d["zmin"])
    try:
        ctx.build_potential(d["xyz"], d["type"], d["npot"], d["vr"], d["pot"], d["dr"],
                            expot=d["expot"])
        dE, _, _ = ctx.energy_range()
        zn, _ = sep.leja_points(nc)
        an = sep.newton_coeffs(zn, np.linspace(d["Elmin"], d["Elmax"], ms), dE, ctx.vmin,
                               (nc / (2.5 * dE)) ** 2)
        print("case trim", flush=True)
        a, _ = ctx.filter(psi0, an, zn, ctx.vmin, dE, sep.SEP_PREC_FP32)
        ctx.trim()
        ctx.trim()  # twice: nothing left to free
        b, _ = ctx.filter(psi0, an, zn, ctx.vmin, dE, sep.SEP_PREC_FP32)
        assert np.array_equal(a, b)
    finally:
        ctx.close()
