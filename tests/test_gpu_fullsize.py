"""The full-size workload (CdSe 4.0 nm dot, 128^3 grid, 16 random states x
32 targets, Newton length 4096 = 512 filtered states) through both
pipelines on the device, checked by size-independent properties: the FP64
pipeline keeps fewer states than it filtered and its Ritz states converge
(sigma < 1e-4) near the band edge; the FP32 pipeline (complex64 filter,
FP64-Gram ortho, FP64 Rayleigh-Ritz, FP64 correction passes on the pairs of
the band-edge window it finds itself, semi_empirical_pseudopotentials_amd/
levels.py) reproduces every such FP64 level within 0.5 eV of the lowest to
1 meV one to one, multiplet by multiplet, and its Ritz values are ascending
with finite variances."""
import numpy as np
import pytest

from semi_empirical_pseudopotentials_amd import levels
from tests.deckutil import compare_levels

pytestmark = pytest.mark.gpu

MEV_HA = 1.0e-3 / 27.211386245988


def test_fullsize_fp32_pipeline_within_1mev(sep):
    import tempfile
    import torch
    from semi_empirical_pseudopotentials_amd import decks
    from semi_empirical_pseudopotentials_amd._lib import read_deck
    n, ns, ms, nc = 128, 16, 32, 4096
    syms, xyz = decks.cdse_dot(4.0)
    tmp = tempfile.mkdtemp(prefix="sep_full_")
    c, h = decks.WINDOW_CENTER_HA, decks.WINDOW_HALF_HA
    decks.write_deck(tmp, syms, xyz, n, n, n, ms, ns, nc, c - h, c + h)
    d = read_deck(tmp)
    N = n ** 3
    psi0 = np.empty((ns, N))
    idum, rng = -1002, None
    for j in range(ns):
        psi0[j], idum, rng = sep.init_psi(idum, N, d["dv"], rng)
    psi0_dev = torch.from_numpy(psi0).to("cuda:0")
    del psi0
    ctx = sep.Context(n, n, n, d["dx"], d["dy"], d["dz"], d["xmin"], d["ymin"], d["zmin"])
    try:
        ctx.build_potential(d["xyz"], d["type"], d["npot"], d["vr"], d["pot"], d["dr"],
                            expot=d["expot"])
        dE, _, _ = ctx.energy_range()
        zn, _ = sep.leja_points(nc)
        an = sep.newton_coeffs(zn, np.linspace(d["Elmin"], d["Elmax"], ms), dE, ctx.vmin,
                               (nc / (2.5 * dE)) ** 2)
        print("case fullsize fp64 pipeline", flush=True)
        ctx.filter_dev(psi0_dev.data_ptr(), ns, an, zn, ctx.vmin, dE, sep.SEP_PREC_FP64)
        k64, S64 = ctx.ortho(ns * ms)
        e64 = ctx.rayleigh_ritz(k64)
        s64 = ctx.sigma(k64)
        print("case fullsize fp32 pipeline", flush=True)
        ctx.filter_dev(psi0_dev.data_ptr(), ns, an, zn, ctx.vmin, dE, sep.SEP_PREC_FP32)
        run = levels.fp32_pipeline(ctx, ns * ms)
        k32, e32, s32 = run["kept"], run["ev_w"], run["sg_w"]
    finally:
        ctx.close()
    assert 0 < k64 < ns * ms and S64[k64 - 1] / S64[0] >= 1e-10
    assert np.all(np.diff(e64) >= 0) and np.all(np.diff(e32) >= 0)
    assert np.all(np.isfinite(s32)) and k32 >= k64
    conv = s64 < 1e-4
    assert conv.sum() >= 20
    e0 = e64[conv].min()
    # the FP32 pipeline found the band edge on its own, within 1 meV of FP64's
    assert abs(run["window"][2] - e0) < MEV_HA
    # one to one over the converged multiplets of the edge band (tests/deckutil.py)
    nc, nl, d = compare_levels(e64, s64, e32, s32, e0, e0 + decks.EDGE_BAND_HA, MEV_HA)
    print("fullsize kept", k64, k32, "edge clusters", nc, "levels", nl, "max |dE| meV",
          d / MEV_HA, flush=True)
    assert nl >= 20


def test_fullsize_filter_vs_oracle(sep, oracle):
    """At the full 128^3 grid of the 4.0 nm dot (1667 atoms): the device
    V(r) equals the oracle's to 1e-13 of its range, and a short filter (2
    random states, 4 targets, 48 Newton steps) of both precisions equals the
    oracle's: FP64 to 1e-9, complex64 to 2e-4 relative."""
    import tempfile
    from semi_empirical_pseudopotentials_amd import decks
    from semi_empirical_pseudopotentials_amd._lib import read_deck
    from tests.deckutil import context_for, device_potential, oracle_potential
    n, ns, ms, nc = 128, 2, 4, 48
    syms, xyz = decks.cdse_dot(4.0)
    tmp = tempfile.mkdtemp(prefix="sep_full_o_")
    c, h = decks.WINDOW_CENTER_HA, decks.WINDOW_HALF_HA
    decks.write_deck(tmp, syms, xyz, n, n, n, ms, ns, nc, c - h, c + h)
    d = read_deck(tmp)
    grid = (n, n, n)
    print("case fullsize oracle V", flush=True)
    V = oracle_potential(oracle, d)
    ksq = oracle.ksqr(n, n, n, d["dx"], d["dy"], d["dz"])
    r = oracle.Rng()
    idc = np.array([-1001], dtype=np.int64)
    psi0 = np.array([r.init_psi(idc, V.size, d["dv"]) for _ in range(ns)])
    with context_for(sep, d) as ctx:
        Vd = device_potential(ctx, d)  # built on the device, returned to the host
        ctx.set_potential(V)
        dE, _, _ = ctx.energy_range()
        zn, _ = sep.leja_points(nc)
        an = sep.newton_coeffs(zn, np.linspace(d["Elmin"], d["Elmax"], ms), dE, V.min(),
                               (nc / (2.5 * dE)) ** 2)
        print("case fullsize filters", flush=True)
        g64, _ = ctx.filter(psi0, an, zn, V.min(), dE, sep.SEP_PREC_FP64)
        g32, _ = ctx.filter(psi0, an, zn, V.min(), dE, sep.SEP_PREC_FP32)
    print("fullsize V max |diff|", np.abs(Vd - V).max(), "range", V.max() - V.min(), flush=True)
    assert np.abs(Vd - V).max() <= 1e-13 * (V.max() - V.min())
    print("case fullsize oracle filter", flush=True)
    ref = np.concatenate([oracle.normalize_all(
        oracle.filter_state(grid, ksq, V, an, zn, 4.0 / dE, V.min(), p), d["dv"]) for p in psi0])
    scale = np.abs(ref).max()
    e64 = np.abs(g64 - ref).max() / scale
    e32 = np.abs(g32 - ref).max() / scale
    print("fullsize filter rel err fp64", e64, "fp32", e32, flush=True)
    # measured 3e-14 (FP64) and 3.4e-6 (complex64, 48 steps): bounds ~15-30x
    assert e64 < 1e-12
    assert e32 < 5e-5


def test_config5_hf_adjoint_vs_fd(sep):
    """The differentiable configuration's grid (3.0 nm dot, 96^3): the
    Hellmann-Feynman adjoint's dR of dv sum_g rho V equals central finite
    differences (h = 1e-6 bohr) of the device V(r) builder to 1e-5 of the
    largest component (plus the differences' own rounding, which matters for
    the off-diagonal density of two random states, whose derivatives nearly
    cancel), for a diagonal and an off-diagonal density."""
    import tempfile
    from semi_empirical_pseudopotentials_amd import decks
    from semi_empirical_pseudopotentials_amd._lib import read_deck
    from tests.deckutil import context_for
    n = 96
    syms, xyz = decks.cdse_dot(3.0)
    tmp = tempfile.mkdtemp(prefix="sep_c5_")
    c, h = decks.WINDOW_CENTER_HA, decks.WINDOW_HALF_HA
    decks.write_deck(tmp, syms, xyz, n, n, n, 8, 8, 64, c - h, c + h)
    d = read_deck(tmp)
    rng = np.random.default_rng(11)
    psi = rng.standard_normal((2, n ** 3))
    psi /= np.sqrt((psi * psi).sum(1) * d["dv"])[:, None]
    print("case config5 adjoint", d["natom"], "atoms", flush=True)
    with context_for(sep, d) as ctx:
        ctx.states_upload(psi)
        # expot depends on |r| alone, so it drops out of d/dR
        dR, _, _ = ctx.hf_adjoint([[0, 0], [0, 1]], d["xyz"], d["type"], d["npot"], d["vr"],
                                  d["pot"], d["dr"])
        hstep = 1e-6
        for p, rho in ((0, psi[0] * psi[0]), (1, psi[0] * psi[1])):
            big = np.abs(dR[p]).max()
            for a in (0, d["natom"] // 2, d["natom"] - 1):
                for k in range(3):
                    q = []
                    for sgn in (1, -1):
                        x = d["xyz"].copy()
                        x[a, k] += sgn * hstep
                        V = ctx.build_potential(x, d["type"], d["npot"], d["vr"], d["pot"],
                                                d["dr"], expot=d["expot"])
                        q.append(np.dot(rho, V) * d["dv"])
                    fd = (q[0] - q[1]) / (2 * hstep)
                    # the differences' own rounding: ~eps sum |rho V| dv / h
                    noise = 1e-15 * np.dot(np.abs(rho), np.abs(V)) * d["dv"] / hstep
                    assert abs(fd - dR[p, a, k]) < 1e-5 * big + noise, (p, a, k, fd, dR[p, a, k])


def test_physical_cdse_conduction_edge_fp32_vs_fp64(sep, tmp_path):
    """A bound workload at full size: the 4.0 nm CdSe dot (1667 atoms,
    128^3) on the reference's own pseudopotential tables (potCd.par and
    potSe.par of filterStrainCube/test/In13_P16, the analytic P1/P2
    passivation of filter/read.c) has a gap (~2.3 eV: valence top near -0.234
    Ha, conduction edge near -0.151 Ha).  Window [-0.26, -0.12] Ha, 16 random
    states x 16 targets, nc 4096.  The complex64 pipeline locates the
    conduction edge from its own Ritz values (the largest gap among pairs
    with sigma < 1e-2), refines the pairs from 0.1 eV below it to 1.1 eV
    above it in place (10 FP64 passes), and every conduction level the FP64
    pipeline converges (sigma < 1e-4) within 1 eV of its edge is matched one
    to one, multiplet by multiplet, within 1 meV.  (The valence top is dense
    -- about 100 levels within 26 mHa -- and converges only to sigma ~1e-4 in
    FP64 at this shape, so no one-to-one check is made there.)"""
    import os
    import shutil
    from semi_empirical_pseudopotentials_amd import decks
    from semi_empirical_pseudopotentials_amd._lib import read_deck
    from tests.conftest import GOLDEN
    from tests.deckutil import context_for, device_potential
    ev_ha = 1.0 / 27.211386245988
    mev = 1e-3 * ev_ha
    n, ns, ms, nc, lo, hi = 128, 16, 16, 4096, -0.26, -0.12
    path = str(tmp_path / "cdse_phys")
    syms, xyz = decks.cdse_dot(4.0)
    decks.write_deck(path, syms, xyz, n, n, n, ms, ns, nc, lo, hi, pots={})
    for s in ("Cd", "Se"):
        shutil.copy(os.path.join(GOLDEN, "In13_P16", f"pot{s}.par"), path)
    d = read_deck(path)
    print("case physical CdSe 4.0 nm 128^3", d["natom"], "atoms", flush=True)
    with context_for(sep, d) as ctx:
        device_potential(ctx, d)
        dE, _, _ = ctx.energy_range()
        zn, _ = sep.leja_points(nc)
        an = sep.newton_coeffs(zn, np.linspace(lo, hi, ms), dE, ctx.vmin, (nc / (2.5 * dE)) ** 2)
        N = n ** 3
        psi0 = np.empty((ns, N))
        idum, rng = -1001, None
        for j in range(ns):
            psi0[j], idum, rng = sep.init_psi(idum, N, d["dv"], rng)
        ctx.filter(psi0, an, zn, ctx.vmin, dE, sep.SEP_PREC_FP64, want_states=False)
        k64, _ = ctx.ortho(ns * ms)
        e64 = ctx.rayleigh_ritz(k64)
        s64 = ctx.sigma(k64)
        print("case physical CdSe complex64 pipeline", flush=True)
        ctx.filter(psi0, an, zn, ctx.vmin, dE, sep.SEP_PREC_FP32, want_states=False)
        k32, _ = ctx.ortho_fp32(ns * ms)
        ev = ctx.rayleigh_ritz(k32)
        sg = ctx.sigma(k32)
        ok = np.sort(ev[(sg < 1e-2) & (ev >= lo) & (ev <= hi)])
        cbm = ok[np.argmax(np.diff(ok)) + 1]
        ev32, sg32 = ctx.refine_window(ev, sg, cbm - 0.1 * ev_ha, cbm + 1.1 * ev_ha, passes=10,
                                       full=True)
    ok64 = np.sort(e64[(s64 < 1e-2) & (e64 >= lo) & (e64 <= hi)])
    cbm64 = ok64[np.argmax(np.diff(ok64)) + 1]
    print(f"conduction edge FP64 {cbm64:.6f} complex64 {cbm:.6f}; gap "
          f"{(cbm64 - ok64[np.argmax(np.diff(ok64))]) / ev_ha:.3f} eV", flush=True)
    assert abs(cbm - cbm64) < 1e-3
    nclu, nl, dmax = compare_levels(e64, s64, ev32, sg32, cbm64 - mev, cbm64 + ev_ha, mev)
    print(f"conduction band: {nl} levels in {nclu} multiplets one to one, max |dE| "
          f"{dmax / mev:.4f} meV", flush=True)
    assert nl >= 8
