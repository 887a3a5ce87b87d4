"""Band-edge levels of the FP32 pipeline, and the one-to-one comparison of
two pipelines' levels.

The FP32 pipeline (no counterpart in the reference, whose filter/main.c:108-
142 runs FP64 throughout): complex64 filter -> sep_ortho_fp32 -> FP64
Rayleigh-Ritz -> sigma, then FP64 correction passes (sep_refine_window:
Rayleigh-Ritz on [Psi, P R, Psi_prev], P the kinetic preconditioner) on the
Ritz pairs of the band-edge window only.  The complex64 filter leaves the
Ritz vectors with residuals near 1e-3 Ha; one pass over all pairs
(sep_refine) brings most edge levels within 1 meV of FP64 but leaves
members of degenerate multiplets up to 2 meV off, and more passes over all
512 pairs cost more than the FP64 pipeline saves.  Ten passes over the ~65
pairs of the edge window bring every edge level within 0.7 meV (128^3, 512
filtered states).
"""
import numpy as np

HARTREE_EV = 27.211386245988
EV = 1.0 / HARTREE_EV
MEV = 1.0e-3 / HARTREE_EV


def compare_levels(e_ref, s_ref, e_got, s_got, lo, hi, tol, conv=1e-4, loose=1e-2):
    """One-to-one comparison of two pipelines' levels in [lo, hi].

    The reference levels with sigma < loose are grouped into clusters (runs
    closer than 2 tol: degenerate multiplets); a cluster is checked when all
    its members have sigma < conv and lie in [lo, hi].  For each such cluster
    the other pipeline must have exactly as many levels with sigma < loose
    within tol of the cluster's span, and the sorted members must agree one
    to one within tol -- a lost level, two merged levels or an extra level
    in a multiplet fail (AssertionError naming the cluster).  Returns
    (clusters checked, levels checked, max |dE|)."""
    e_ref, s_ref = np.asarray(e_ref), np.asarray(s_ref)
    e_got, s_got = np.asarray(e_got), np.asarray(s_got)
    order = np.argsort(e_ref)
    e_ref, s_ref = e_ref[order], s_ref[order]
    sel = s_ref < loose
    r, rs = e_ref[sel], s_ref[sel]
    g = np.sort(e_got[s_got < loose])
    clusters, cur = [], ([0] if r.size else [])
    for i in range(1, r.size):
        if r[i] - r[i - 1] < 2 * tol:
            cur.append(i)
        else:
            clusters.append(cur)
            cur = [i]
    if cur:
        clusters.append(cur)
    nc = nl = 0
    worst = 0.0
    for cl in clusters:
        ec = r[cl]
        if not (np.all(rs[cl] < conv) and ec[0] >= lo and ec[-1] <= hi):
            continue
        mine = g[(g >= ec[0] - tol) & (g <= ec[-1] + tol)]
        if mine.size != ec.size:
            near = (e_got >= ec[0] - 4 * tol) & (e_got <= ec[-1] + 4 * tol)
            raise AssertionError(
                f"cluster at {ec[0]:.8f} Ha: {ec.size} reference levels {ec}, {mine.size} "
                f"matched; other pipeline nearby: {e_got[near]} sigma {s_got[near]}")
        d = np.abs(mine - ec).max()
        if d > tol:
            raise AssertionError(f"cluster at {ec[0]:.8f} Ha: |dE| {d:.3e} > {tol:.3e}")
        worst = max(worst, d)
        nc += 1
        nl += ec.size
    return nc, nl, worst


def edge_window(ev, sg, band=0.5 * EV, below=0.1 * EV, above=0.1 * EV, sig_edge=1e-2):
    """[e0 - below, e0 + band + above] around the band edge e0, the lowest
    Ritz value with sigma < sig_edge (unrefined complex64 Ritz pairs of real
    levels sit near 1e-3, those of the noise directions far above)."""
    ok = np.asarray(sg) < sig_edge
    e0 = float(np.min(np.asarray(ev)[ok])) if ok.any() else float(np.min(ev))
    return e0 - below, e0 + band + above, e0


def fp32_pipeline(ctx, m, passes=10, band=0.5 * EV):
    """FP32 ortho + FP64 Rayleigh-Ritz + sigma on the m resident filtered
    states of a complex64 filter, then `passes` FP64 correction passes on the
    pairs of the band-edge window (refined in place among the resident
    pairs).  Returns a dict: kept, ev / sg of all Ritz pairs before the
    passes, window (lo, hi, e0), ev_w / sg_w of the refined pairs in the
    window."""
    kept, _ = ctx.ortho_fp32(m)
    ev = ctx.rayleigh_ritz(kept)
    sg = ctx.sigma(kept)
    lo, hi, e0 = edge_window(ev, sg, band)
    ev_w, sg_w = ctx.refine_window(ev, sg, lo, hi, passes=passes)
    return {"kept": kept, "ev": ev, "sg": sg, "window": (lo, hi, e0), "ev_w": ev_w, "sg_w": sg_w}
