"""Batched uneven grids (sep_filter_batch): several nanocrystals with
different grids filtered in one call, interleaved on their streams or as one
captured HIP graph per grid, give exactly the states of one sep_filter_dev
call per grid (same kernels in the same order per grid)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dots(sep):
    import tempfile
    import torch
    from semi_empirical_pseudopotentials_amd import decks
    from semi_empirical_pseudopotentials_amd._lib import read_deck
    out = []
    for k, (dnm, n, ns) in enumerate(((1.2, 24, 3), (1.5, 32, 2), (1.8, 48, 4))):
        tmp = tempfile.mkdtemp(prefix="sep_batch_")
        syms, xyz = decks.cdse_dot(dnm)
        c, h = decks.WINDOW_CENTER_HA, decks.WINDOW_HALF_HA
        decks.write_deck(tmp, syms, xyz, n, n, n, 4, ns, 64 + 16 * k, c - h, c + h)
        d = read_deck(tmp)
        ctx = sep.Context(n, n, n, d["dx"], d["dy"], d["dz"], d["xmin"], d["ymin"], d["zmin"])
        ctx.build_potential(d["xyz"], d["type"], d["npot"], d["vr"], d["pot"], d["dr"],
                            expot=d["expot"])
        dE, _, _ = ctx.energy_range()
        nc = 64 + 16 * k
        zn, _ = sep.leja_points(nc)
        an = sep.newton_coeffs(zn, np.linspace(d["Elmin"], d["Elmax"], 4), dE, ctx.vmin,
                               (nc / (2.5 * dE)) ** 2)
        N = n ** 3
        psi0 = np.empty((ns, N))
        idum, rng = -1001 - k, None
        for j in range(ns):
            psi0[j], idum, rng = sep.init_psi(idum, N, d["dv"], rng)
        out.append(dict(ctx=ctx, ns=ns, an=an, zn=zn, dE=dE, vmin=ctx.vmin,
                        psi0=torch.from_numpy(psi0).to("cuda:0")))
    yield out
    for o in out:
        o["ctx"].close()


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
@pytest.mark.parametrize("mode", ["interleave", "graph"])
def test_batch_equals_single(sep, dots, prec, mode):
    P = sep.SEP_PREC_FP32 if prec == "fp32" else sep.SEP_PREC_FP64
    M = sep.SEP_BATCH_INTERLEAVE if mode == "interleave" else sep.SEP_BATCH_GRAPH
    print("case batch", prec, mode, flush=True)
    single = []
    for o in dots:
        o["ctx"].filter_dev(o["psi0"].data_ptr(), o["ns"], o["an"], o["zn"], o["vmin"], o["dE"], P)
        single.append(o["ctx"].states_download(o["ns"] * o["an"].shape[0]))
    sep.filter_batch([o["ctx"] for o in dots], [o["psi0"].data_ptr() for o in dots],
                     [o["ns"] for o in dots], [o["an"] for o in dots], [o["zn"] for o in dots],
                     [o["vmin"] for o in dots], [o["dE"] for o in dots], P, M)
    for o, ref in zip(dots, single):
        got = o["ctx"].states_download(o["ns"] * o["an"].shape[0])
        assert np.array_equal(got, ref)


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_autotune_choice_keeps_results(sep, dots, prec):
    """The autotuner only picks between two correct kernel families: H.psi after
    it agrees with H.psi before it (plane vs row + line kernels)."""
    P = sep.SEP_PREC_FP32 if prec == "fp32" else sep.SEP_PREC_FP64
    o = dots[2]
    ctx = o["ctx"]
    N = ctx.N
    rng = np.random.default_rng(7)
    x = rng.standard_normal((2, N)) + 1j * rng.standard_normal((2, N))
    before = ctx.apply_h(x) if P == sep.SEP_PREC_FP64 else None
    tp, tl = ctx.autotune(P, 2)
    print("autotune", prec, tp, tl, flush=True)
    assert tl > 0.0 and (tp > 0.0 or tp == -1.0)
    if before is not None:
        after = ctx.apply_h(x)
        assert np.abs(after - before).max() <= 1e-12 * np.abs(before).max()
