"""The filter loop's fused kernels (pass C of Newton step j with pass A of
step j+1: k_plane_bwd_fwd_w, and k_rows_bwd_fwd for the row + line family)
against the same loop with the passes as separate kernels (test hook
SEP_SEPARATE_PASSES), for odd and even state counts, on plane lengths with
256-, 512- and 1024-thread plane kernels and on the line family.
FP64 agrees to 1e-13.  FP32 agrees to rounding only: the compiler contracts
the complex products of the two kernels into FMAs differently, and the
Newton recurrence carries those single-rounding differences through the
steps (measured 3-6e-6 relative after 48 steps); the fused FP32 states are
as close to the FP64 filter's as the separate kernels' are."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup(sep, n, dnm, ns, nc, ms):
    import tempfile
    from semi_empirical_pseudopotentials_amd import decks
    from semi_empirical_pseudopotentials_amd._lib import read_deck
    tmp = tempfile.mkdtemp(prefix="sep_fused_")
    syms, xyz = decks.cdse_dot(dnm)
    c, h = decks.WINDOW_CENTER_HA, decks.WINDOW_HALF_HA
    decks.write_deck(tmp, syms, xyz, n, n, n, ms, ns, nc, c - h, c + h)
    d = read_deck(tmp)
    N = n ** 3
    psi0 = np.empty((ns, N))
    idum, rng = -1001, None
    for j in range(ns):
        psi0[j], idum, rng = sep.init_psi(idum, N, d["dv"], rng)
    return d, psi0


def _run(sep, d, psi0, nc, ms, prec):
    n = d["nx"]
    ctx = sep.Context(n, n, n, d["dx"], d["dy"], d["dz"], d["xmin"], d["ymin"], d["zmin"])
    try:
        ctx.build_potential(d["xyz"], d["type"], d["npot"], d["vr"], d["pot"], d["dr"],
                            expot=d["expot"])
        dE, _, _ = ctx.energy_range()
        zn, _ = sep.leja_points(nc)
        an = sep.newton_coeffs(zn, np.linspace(d["Elmin"], d["Elmax"], ms), dE, ctx.vmin,
                               (nc / (2.5 * dE)) ** 2)
        states, _ = ctx.filter(psi0, an, zn, ctx.vmin, dE, prec)
        kms = ctx.bench_hpsi(prec, 2, ms, 1)
        return states, kms
    finally:
        ctx.close()


@pytest.mark.parametrize("family", ["plane", "lines"])
@pytest.mark.parametrize("prec", ["fp32", "fp64"])
@pytest.mark.parametrize("n,ns", [(32, 3), (64, 4), (96, 2), (128, 1)])
def test_fused_equals_separate(sep, monkeypatch, family, prec, n, ns):
    if family == "lines":  # k_rows_bwd_fwd (FP64 at 128 takes it without the hook)
        monkeypatch.setenv("SEP_FORCE_LINE_KERNELS", "1")
    P = sep.SEP_PREC_FP32 if prec == "fp32" else sep.SEP_PREC_FP64
    nc, ms = 48, 3
    d, psi0 = _setup(sep, n, 1.2 + n / 64.0, ns, nc, ms)
    print("case fused", family, prec, n, ns, flush=True)
    fused, kms = _run(sep, d, psi0, nc, ms, P)
    assert kms[5] > 0.0  # the fused kernel exists for this grid
    monkeypatch.setenv("SEP_SEPARATE_PASSES", "1")
    sepr, kms2 = _run(sep, d, psi0, nc, ms, P)
    assert kms2[5] > 0.0  # the hook only changes the filter loop
    scale = np.abs(sepr).max()
    err = np.abs(fused - sepr).max() / scale
    print("max rel diff", err, "bitwise", np.array_equal(fused, sepr), flush=True)
    if prec == "fp64":
        assert err < 1e-13
        return
    assert err < 2e-5
    ref, _ = _run(sep, d, psi0, nc, ms, sep.SEP_PREC_FP64)
    e_f = np.abs(fused - ref).max() / scale
    e_s = np.abs(sepr - ref).max() / scale
    print("vs fp64: fused", e_f, "separate", e_s, flush=True)
    assert e_f <= 1.5 * e_s + 1e-7
