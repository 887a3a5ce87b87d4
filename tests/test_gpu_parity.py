"""End-to-end parity of the FP64 GPU pipeline (the `filter` drop-in) with the
reference binary's own outputs and with the CPU oracle, eigenvalues and Ritz
vectors, and the FP32 pipeline's per-eigenvalue error against the FP64 oracle.

* In13_P16 (filterStrainCube/test/In13_P16, 36^3, 16 x 10 states, nc 1024) at
  the seed of the run that wrote the reference's eval.dat and
  eval-filt-*.dat (recovered, see tests/test_oracle_deck.py): filtered-state
  energies against the reference files and the oracle, eval.dat row by row,
  and Ritz-vector overlaps against the oracle's psi.dat (degenerate levels
  compared as subspaces through principal angles).
* Config C1 at full size (2.0 nm CdSe, 64^3, 8 targets x 4 states, nc 1024)
  against the committed oracle fixture tests/golden/c1/c1_oracle.npz
  (eigenvalues, filtered energies, sketched Ritz-vector overlaps).
* Config C3: the same workload through the FP32 pipeline, error per
  eigenvalue against the C1 FP64 oracle.
"""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

from tests.conftest import GOLDEN
from tests.deckutil import read_eval

pytestmark = pytest.mark.gpu

sys.path.insert(0, os.path.join(GOLDEN, "c1"))
import make_fixture as c1fix  # noqa: E402

IN13 = os.path.join(GOLDEN, "In13_P16")
IN13_SEED = -822262754
HA_MEV = 27.211386245988e3


def _run_filter(deck, out, *args):
    from semi_empirical_pseudopotentials_amd._lib import FILTER_BIN
    shutil.copytree(deck, out, dirs_exist_ok=True)
    r = subprocess.run([FILTER_BIN, "--dir", out, *args], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout


def clusters(e, tol):
    """Runs of consecutive (ascending) levels closer than tol."""
    out, cur = [], [0]
    for i in range(1, len(e)):
        if e[i] - e[i - 1] < tol:
            cur.append(i)
        else:
            out.append(cur)
            cur = [i]
    out.append(cur)
    return out


def overlap_defect(Ua, Ub):
    """1 - cos(largest principal angle) between the row spans of Ua, Ub
    (rows orthonormal under the plain dot product): for one vector each,
    1 - |<a|b>|."""
    s = np.linalg.svd(Ua @ Ub.T, compute_uv=False)
    return 1.0 - s.min()


def _unit(psi, dv):
    return psi * np.sqrt(dv)


def test_in13_reference_and_oracle(sep, oracle, tmp_path):
    """Filtered-state energies: GPU vs the reference's eval-filt files and vs
    the oracle to 5e-9 Ha -- they move with the spectral bound dE, which the
    power iteration (filter/energy.c:53-83, stopped at a 1e-4 relative change)
    fixes only to the rounding of its FFTs, ~1e-10 relative here, and a
    1e-10 shift of dE moves the filter centres by ~1e-9 Ha.  The same shift
    perturbs the filtered subspace itself at ~1e-9, which the Ritz values of
    converged levels feel only at second order.  eval.dat: 158 kept; levels
    with sigma < 1e-4 within 1e-9 Ha of the oracle and 1e-8 Ha of the
    reference (whose per-atom scales carry 8 decimals), levels with sigma <
    1e-2 within 1e-8 Ha of both; the Ritz vectors of the sigma < 1e-4 levels
    overlap the oracle's to 1 - 1e-8 (clusters closer than 1e-6 Ha compared
    as subspaces)."""
    from tests.test_oracle_deck import read_filt_fixtures
    print("case In13_P16 seed", IN13_SEED, flush=True)
    gout = str(tmp_path / "gpu")
    log = _run_filter(IN13, gout, "--seed", str(IN13_SEED))
    assert "dE = 13.3627" in log and "cutoff is 158" in log
    print("case In13_P16 oracle", flush=True)
    oout = str(tmp_path / "cpu")
    s = oracle.run(IN13, oout, IN13_SEED, 0)
    assert s["cutoff"] == 158
    ref = read_filt_fixtures(IN13)
    w_ref = w_or = 0.0
    for jns, (e_ref, _) in ref.items():
        g = np.loadtxt(os.path.join(gout, f"eval-filt-0-{jns}.dat"), ndmin=2)[:, 1]
        o = np.loadtxt(os.path.join(oout, f"eval-filt-0-{jns}.dat"), ndmin=2)[:, 1]
        w_ref = max(w_ref, np.abs(g - e_ref).max())
        w_or = max(w_or, np.abs(g - o).max())
    print("eval-filt: vs reference", w_ref, "vs oracle", w_or, flush=True)
    assert w_ref < 5e-9 and w_or < 5e-9
    ref_e, ref_s = read_eval(os.path.join(IN13, "eval.dat"))
    ge, gs = read_eval(os.path.join(gout, "eval.dat"))
    oe, os_ = read_eval(os.path.join(oout, "eval.dat"))
    assert ge.size == oe.size == ref_e.size == 158
    conv = (ref_s < 1e-2) & (os_ < 1e-2)
    tight = (ref_s < 1e-4) & (os_ < 1e-4)
    d_ref, d_or = np.abs(ge - ref_e), np.abs(ge - oe)
    print("eval.dat sigma < 1e-4:", tight.sum(), "levels, vs reference", d_ref[tight].max(),
          "vs oracle", d_or[tight].max(), flush=True)
    print("eval.dat sigma < 1e-2:", conv.sum(), "levels, vs reference", d_ref[conv].max(),
          "vs oracle", d_or[conv].max(), "; rest vs oracle", d_or[~conv].max(), flush=True)
    assert tight.sum() >= 50 and conv.sum() >= 80
    assert d_or[tight].max() < 1e-9 and d_ref[tight].max() < 1e-8
    assert d_or[conv].max() < 1e-8 and d_ref[conv].max() < 1e-8
    assert d_or[~conv].max() < 1e-4
    N, dv = 36 ** 3, s["dv"]
    Ug = _unit(np.fromfile(os.path.join(gout, "psi.dat")).reshape(158, N), dv)
    Uo = _unit(np.fromfile(os.path.join(oout, "psi.dat")).reshape(158, N), dv)
    worst, ncl = 0.0, 0
    for c in clusters(oe, 1e-6):
        if not all(tight[i] for i in c):
            continue
        worst = max(worst, overlap_defect(Ug[c], Uo[c]))
        ncl += 1
    print("Ritz vectors: clusters", ncl, "max 1 - overlap", worst, flush=True)
    assert ncl >= 25
    assert worst < 1e-8


@pytest.fixture(scope="module")
def c1(tmp_path_factory):
    ref = np.load(os.path.join(GOLDEN, "c1", "c1_oracle.npz"))
    deck = c1fix.write_c1_deck(str(tmp_path_factory.mktemp("c1_deck")))
    assert c1fix.deck_digest(deck) == str(ref["digest"]), "C1 deck generator changed"
    return deck, ref


def test_c1_fp64_full_size(sep, c1, tmp_path):
    """Config C1 at its full size.  32 filtered states of the gapless synthetic
    spectrum, none converged (sigma 2e-4 .. 1.6 Ha): the Ritz values are
    fixed by FP64 arithmetic only to about 1e-3 sigma^2 Ha, because the
    subspace's weakest directions (S_31/S_0 = 2.4e-6) amplify rounding-level
    differences of the Newton recurrence -- the oracle built with FMA
    contraction (eval_fma in the fixture) differs from its default build by
    exactly that much (3.6e-11 Ha at the lowest level, 2.2e-3 Ha at the
    highest).  Asserted: kept count; filtered energies to 5e-9 Ha (they move
    with dE, see the In13_P16 test); every Ritz value within 1e-9 Ha + 3x the
    oracle's own rounding spread of that level; the levels the oracle fixes
    to 1e-10 Ha (sigma < 3e-4) within 1e-9 Ha; their Ritz vectors within a
    sketched 1 - overlap of 1e-8."""
    deck, ref = c1
    print("case C1 fp64", flush=True)
    out = str(tmp_path / "gpu")
    log = _run_filter(deck, out, "--seed", str(c1fix.C1["seed"]))
    assert "cutoff is 32" in log
    ge, gs = read_eval(os.path.join(out, "eval.dat"))
    kept = int(ref["kept"])
    assert ge.size == kept
    wf = max(np.abs(np.loadtxt(os.path.join(out, f"eval-filt-0-{j}.dat"), ndmin=2)[:, 1]
                    - ref["filt"][j]).max() for j in range(c1fix.C1["ns"]))
    e_or = ref["eval"]
    spread = np.abs(ref["eval_fma"] - e_or)
    de = np.abs(ge - e_or)
    for i in range(kept):
        print(f"  C1 level {i}: gpu {ge[i]:.12f} oracle {e_or[i]:.12f} |d| {de[i]:.2e} "
              f"oracle rounding spread {spread[i]:.2e} sigma {gs[i]:.3e} / {ref['sigma'][i]:.3e}",
              flush=True)
    print("C1 filtered energies max |d|", wf, flush=True)
    assert wf < 5e-9
    assert np.all(de <= 1e-9 + 3.0 * spread)
    fixed = spread < 1e-10
    assert fixed.sum() >= 2 and de[fixed].max() < 1e-9
    N = c1fix.C1["n"] ** 3
    psi = np.fromfile(os.path.join(out, "psi.dat")).reshape(kept, N)
    Y = c1fix.sketch(psi, float(ref["dv"]), c1fix.sketch_matrix(N))
    Yo = ref["sketch"]
    est = []
    for c in clusters(e_or, 1e-6):
        if not all(fixed[i] for i in c):
            continue
        if len(c) == 1:
            i = c[0]
            dmin = min(np.sum((Y[i] - Yo[i]) ** 2), np.sum((Y[i] + Yo[i]) ** 2))
            est.append(0.5 * dmin)
        else:  # sketched projector distance ~ 2 sum sin^2 of principal angles
            Pd = Y[c].T @ Y[c] - Yo[c].T @ Yo[c]
            est.append(0.25 * np.sum(Pd * Pd))
    print("C1 sketched 1 - overlap of the fixed levels", est, flush=True)
    assert len(est) >= 2 and max(est) < 1e-8


def _fp32_pipeline(sep, deck, ns):
    """The FP32 pipeline through the C ABI; Ritz values before and after the
    correction pass, and the sigmas after it."""
    from semi_empirical_pseudopotentials_amd._lib import read_deck
    d = read_deck(deck)
    n, ms, nc = c1fix.C1["n"], c1fix.C1["ms"], c1fix.C1["nc"]
    with sep.Context(n, n, n, d["dx"], d["dy"], d["dz"], d["xmin"], d["ymin"], d["zmin"]) as ctx:
        ctx.build_potential(d["xyz"], d["type"], d["npot"], d["vr"], d["pot"], d["dr"],
                            expot=d["expot"])
        vmin = ctx.vmin
        dE, _, _ = ctx.energy_range()
        zn, _ = sep.leja_points(nc)
        an = sep.newton_coeffs(zn, np.linspace(d["Elmin"], d["Elmax"], ms), dE, vmin,
                               (nc / (2.5 * dE)) ** 2)
        N = n ** 3
        psi0 = np.empty((ns, N))
        idum, rng = c1fix.C1["seed"], None
        for j in range(ns):
            psi0[j], idum, rng = sep.init_psi(idum, N, d["dv"], rng)
        ctx.filter(psi0, an, zn, vmin, dE, sep.SEP_PREC_FP32, want_states=False)
        k, _ = ctx.ortho_fp32(ns * ms)
        ev_rr = ctx.rayleigh_ritz(k)
        ev_c, sg_c = ctx.refine(ev_rr)
    return k, ev_rr, ev_c, sg_c


def test_c3_fp32_per_eigenvalue(sep, c1, tmp_path):
    """Config C3: the C1 dot through the FP32 pipeline (complex64 filter, FP32
    pipeline ortho, FP64 Rayleigh-Ritz, one FP64 correction pass), error per
    eigenvalue against the FP64 oracle.

    At C1's 4 random states no level is converged, so an FP64 Ritz value is
    itself within only ~sigma^2/gap of an eigenvalue and the complex64
    recurrence's rounding reaches the unconverged Ritz pairs through the same
    weak directions as in the FP64 test, 1e9 times larger: the table is
    printed, not asserted.  The 1 meV bar is asserted on the same dot, grid,
    targets and Newton length with 16 random states, where the band-edge
    levels converge: every oracle level with sigma < 1e-4 within 0.5 eV of
    the lowest is matched, one to one (as many pipeline levels in the window
    as oracle levels, sorted arrays compared element by element), by the
    FP32 pipeline to <= 1 meV (measured <= 0.03 meV).  The pipeline's Ritz
    vectors are less converged than FP64's (sigma 5e-5 .. 1e-3 against
    1e-8 .. 5e-6): printed beside the levels."""
    from semi_empirical_pseudopotentials_amd import decks
    deck, ref = c1
    print("case C3 fp32 pipeline, 4 states (table only)", flush=True)
    k, ev_rr, ev_c, sg_c = _fp32_pipeline(sep, deck, c1fix.C1["ns"])
    e_or = ref["eval"]
    edge = e_or <= e_or[0] + decks.EDGE_BAND_HA
    for i in np.nonzero(edge)[0]:
        print(f"  C3/4 level {i}: oracle {e_or[i]:.10f} Ha sigma {ref['sigma'][i]:.2e}; fp32 RR "
              f"{(ev_rr[i] - e_or[i]) * HA_MEV:+.4f} meV; with correction "
              f"{(ev_c[i] - e_or[i]) * HA_MEV:+.4f} meV (sigma {sg_c[i]:.2e})", flush=True)
    print("case C3 fp32 pipeline, 16 states", flush=True)
    deck3 = c1fix.write_c1_deck(str(tmp_path / "c3"), ns=c1fix.C3_NS)
    assert c1fix.deck_digest(deck3) == str(ref["c3_digest"])
    k3, rr3, ev3, sg3 = _fp32_pipeline(sep, deck3, c1fix.C3_NS)
    e3, s3 = ref["c3_eval"], ref["c3_sigma"]
    conv = s3 < 1e-4
    sel = np.nonzero(conv & (e3 <= e3[conv].min() + decks.EDGE_BAND_HA))[0]
    got = np.sort(ev3)
    assert sel.size >= 5
    errs = []
    for i in sel:
        j = int(np.argmin(np.abs(got - e3[i])))
        errs.append((got[j] - e3[i]) * HA_MEV)
        print(f"  C3/16 level {i}: oracle {e3[i]:.10f} Ha sigma {s3[i]:.2e}; fp32 pipeline "
              f"{errs[-1]:+.5f} meV", flush=True)
    # one to one: the converged FP32 levels in the window pair off with the
    # oracle's without sharing (count equal and sorted arrays aligned)
    lo, hi = e3[sel].min() - 1e-4, e3[sel].max() + 1e-4
    inwin = (ev3 >= lo) & (ev3 <= hi)
    for e, sg in zip(ev3[inwin], sg3[inwin]):
        print(f"  C3/16 fp32 level in window: {e:.10f} Ha sigma {sg:.2e}", flush=True)
    g_win = got[(got >= lo) & (got <= hi)]
    assert g_win.size == sel.size, (g_win.size, sel.size)
    assert np.abs(g_win - np.sort(e3[sel])).max() * HA_MEV <= 1.0
    assert max(abs(e) for e in errs) <= 1.0
