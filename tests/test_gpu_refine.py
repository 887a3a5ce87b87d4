"""The FP32 subspace pipeline (complex64 filter -> FP64-Gram ortho -> FP64
Rayleigh-Ritz -> one FP64 preconditioned correction pass) against the FP64
pipeline on the same seeded states: levels within 0.5 eV of the lowest
converged level agree to 1 meV, and the ortho basis is orthonormal with the
Gram eigenvalues of the double-double FP64 ortho."""
import numpy as np
import pytest

from tests.deckutil import context_for, device_potential

pytestmark = pytest.mark.gpu

MEV = 1.0e-3 / 27.211386245988


@pytest.fixture(scope="module")
def run(sep, tmp_path_factory):
    from semi_empirical_pseudopotentials_amd import decks
    from semi_empirical_pseudopotentials_amd._lib import read_deck
    path = str(tmp_path_factory.mktemp("deck_refine"))
    syms, xyz = decks.cdse_dot(1.5)
    c, h = decks.WINDOW_CENTER_HA, decks.WINDOW_HALF_HA
    decks.write_deck(path, syms, xyz, 48, 48, 48, ms=8, ns=16, nc=1024, elmin=c - h, elmax=c + h)
    d = read_deck(path)
    ctx = context_for(sep, d)
    device_potential(ctx, d)
    vmin = ctx.vmin
    dE, _, _ = ctx.energy_range()
    nc, ms, ns = 1024, 8, 16
    zn, _ = sep.leja_points(nc)
    el = np.linspace(d["Elmin"], d["Elmax"], ms)
    an = sep.newton_coeffs(zn, el, dE, vmin, (nc / (2.5 * dE)) ** 2)
    N = d["nx"] * d["ny"] * d["nz"]
    psi0 = np.empty((ns, N))
    idum, rng = -1003, None
    for j in range(ns):
        psi0[j], idum, rng = sep.init_psi(idum, N, d["dv"], rng)
    out = {"d": d}
    print("case refine fp64 pipeline", flush=True)
    blk64, _ = ctx.filter(psi0, an, zn, vmin, dE, sep.SEP_PREC_FP64)
    out["blk64"] = blk64
    k64, S64 = ctx.ortho(ns * ms)
    ev64 = ctx.rayleigh_ritz(k64)
    sg64 = ctx.sigma(k64)
    out.update(k64=k64, S64=S64, ev64=ev64, sg64=sg64)
    print("case refine fp32 pipeline", flush=True)
    blk32, _ = ctx.filter(psi0, an, zn, vmin, dE, sep.SEP_PREC_FP32)
    k32, S32 = ctx.ortho_fp32(ns * ms)
    Q32 = ctx.states_download(k32)
    ev_rr = ctx.rayleigh_ritz(k32)
    sg_rr = ctx.sigma(k32)
    ev32, sg32 = ctx.refine(ev_rr)
    out.update(k32=k32, S32=S32, Q32=Q32, ev_rr=ev_rr, sg_rr=sg_rr, ev32=ev32, sg32=sg32)
    # window passes in place on the refined pairs: window edges in the widest
    # gaps near a third and a half of the spectrum
    print("case refine window in place", flush=True)
    gaps = np.diff(ev32)
    i = k32 // 3 + int(np.argmax(gaps[k32 // 3:k32 // 3 + 8]))
    j = k32 // 2 + int(np.argmax(gaps[k32 // 2:k32 // 2 + 8]))
    lo, hi = 0.5 * (ev32[i] + ev32[i + 1]), 0.5 * (ev32[j] + ev32[j + 1])
    before = ctx.states_download(k32)
    evw, sgw = ctx.refine_window(ev32, sg32, lo, hi, passes=3, full=True)
    out.update(win=(lo, hi), before=before, after=ctx.states_download(k32), evw=evw, sgw=sgw)
    # the double-double FP64 ortho of the same complex64-filtered block
    ctx.states_upload(blk32)
    out["S64_same"] = ctx.ortho(ns * ms)[1]
    ctx.close()
    return out


def test_ortho_fp32_basis(run):
    d = run["d"]
    # the complex64 filter's rounding noise lifts the small singular values,
    # so the FP32 pipeline can keep more directions than the FP64 one (all
    # 16 x 8 here)
    k32 = run["k32"]
    assert 1 <= k32 <= 16 * 8
    Q = run["Q32"]
    assert np.abs(Q @ Q.T * d["dv"] - np.eye(k32)).max() < 1e-12
    # Gram eigenvalues S^2 as the double-double FP64 ortho's of the same block,
    # to FP64 Gram accuracy (absolute, relative to the largest)
    S32, S64 = run["S32"], run["S64_same"]
    dl = np.abs(S32[:k32] ** 2 - S64[:k32] ** 2).max() / S64[0] ** 2
    print("fp64 Gram eigenvalue error", dl, flush=True)
    assert dl < 1e-13
    assert S32[k32 - 1] / S32[0] >= sep_floor()


def test_refine_lowers_variance(run):
    sg_rr, sg32 = run["sg_rr"], run["sg32"]
    assert np.all(np.isfinite(sg32))
    assert np.median(sg32) <= np.median(sg_rr) * (1 + 1e-9)


def test_refine_window_in_place(run):
    """sep_refine_window refines the pairs in [lo, hi] in place: the other
    resident pairs (states, energies, variances) are untouched, the window's
    variances drop, and the pairs stay ascending in energy."""
    lo, hi = run["win"]
    ev, sg, evw, sgw = run["ev32"], run["sg32"], run["evw"], run["sgw"]
    inw = (ev >= lo) & (ev <= hi)
    assert inw.sum() >= 3 and (~inw).sum() >= 3
    assert np.all(np.diff(evw) >= 0)
    assert np.array_equal(evw[~inw], ev[~inw]) and np.array_equal(sgw[~inw], sg[~inw])
    b, a = run["before"], run["after"]
    assert np.array_equal(a[~inw], b[~inw])
    assert not np.array_equal(a[inw], b[inw])
    print("window pairs", inw.sum(), "median sigma", np.median(sg[inw]), "->",
          np.median(sgw[inw]), flush=True)
    assert np.median(sgw[inw]) < np.median(sg[inw])


def test_fp32_pipeline_edge_levels_within_1mev(run):
    from tests.deckutil import compare_levels
    ev64, sg64, ev32, sg32 = run["ev64"], run["sg64"], run["ev32"], run["sg32"]
    conv = sg64 < 1e-3
    assert conv.sum() >= 3
    e0 = ev64[conv].min()
    nc, nl, d = compare_levels(ev64, sg64, ev32, sg32, e0, e0 + 0.5 / 27.211386245988, MEV,
                               conv=1e-3)
    print("edge clusters", nc, "levels", nl, "max |dE| meV", d / MEV, flush=True)
    assert nl >= 3


def sep_floor():
    return 1.0e-7  # SEP_FP32_SVD_FLOOR in include/sep.h


REFINE_SHIFT = 1.0  # SEP_REFINE_SHIFT in include/sep.h


def _np_refine(oracle, grid, ksq, V, dv, X, e):
    """numpy restatement of the preconditioned correction pass with the
    oracle's H; the preconditioner 1 / (T + shift) with T = N ksq (the oracle's
    kinetic table holds min(k^2/2, Ekinmax) / N) applied by numpy's FFT."""
    nx, ny, nz = grid
    kin = ksq.reshape(nz, ny, nx) * (nx * ny * nz)

    def Hm(Y):
        return oracle.hamiltonian(grid, ksq, V, Y.astype(np.complex128)).real

    def Pm(Y):
        Z = np.fft.fftn(Y.reshape(-1, nz, ny, nx), axes=(1, 2, 3)) / (kin + REFINE_SHIFT)
        return np.fft.ifftn(Z, axes=(1, 2, 3)).real.reshape(Y.shape)

    def basis(A, floor):
        w, W = np.linalg.eigh(A @ A.T)
        w, W = w[::-1], W[:, ::-1]
        s = np.sqrt(np.maximum(w, 0))
        k = int(np.sum(s / s[0] >= floor))
        U = (W[:, :k] / s[:k]).T @ A
        for _ in range(2):
            U = np.linalg.solve(np.linalg.cholesky(U @ U.T), U)
        return U / np.sqrt((U * U).sum(1) * dv)[:, None]

    R = Hm(X) - e[:, None] * X
    R = R / np.sqrt((R * R).sum(1) * dv)[:, None]
    T = Pm(R)
    T = T / np.sqrt((T * T).sum(1) * dv)[:, None]
    U = basis(np.concatenate([X, T]), sep_floor())
    HU = Hm(U)
    P = U @ HU.T * dv
    ev, Q = np.linalg.eigh(0.5 * (P + P.T))
    Y = Q.T @ U
    Y = Y / np.sqrt((Y * Y).sum(1) * dv)[:, None]
    HY = Hm(Y)
    ee = (Y * HY).sum(1) * dv
    sg = np.sqrt(np.abs((HY * HY).sum(1) * dv - ee ** 2))
    idx = np.sort(np.argsort(sg)[:len(e)])
    return ev[idx], sg[idx]


def test_refine_vs_numpy(sep, oracle, small_deck):
    """sep_refine on the oracle's Ritz pairs of a small filtered block equals
    the numpy restatement with the oracle's H (energies 1e-9 Ha, sigma 1e-7)."""
    from semi_empirical_pseudopotentials_amd._lib import read_deck
    from tests.deckutil import oracle_potential
    d = read_deck(small_deck)
    grid = (d["nx"], d["ny"], d["nz"])
    V = oracle_potential(oracle, d)
    ksq = oracle.ksqr(*grid, d["dx"], d["dy"], d["dz"])
    dv = d["dv"]
    dE, _, _ = oracle.energy_range(grid, ksq, V, dv, V.min())
    nc, ms, ns = 96, 4, 3
    zn, _ = oracle.leja_points(nc)
    an = oracle.newton_coeffs(nc, np.linspace(d["Elmin"], d["Elmax"], ms), zn, dE, V.min(),
                              (nc / (2.5 * dE)) ** 2)
    r = oracle.Rng()
    idc = np.array([-1001], dtype=np.int64)
    blk = np.concatenate([oracle.normalize_all(oracle.filter_state(
        grid, ksq, V, an, zn, 4.0 / dE, V.min(), r.init_psi(idc, V.size, dv)), dv)
        for _ in range(ns)])
    U, _, cut = oracle.ortho(blk)
    U = oracle.normalize_all(U[:cut], dv)
    e, X = oracle.hmat(grid, ksq, V, dv, U)
    X = oracle.normalize_all(X, dv)
    e_ref, s_ref = _np_refine(oracle, grid, ksq, V, dv, X, e)
    print("case refine_vs_numpy", cut, flush=True)
    with context_for(sep, d) as ctx:
        ctx.set_potential(V)
        ctx.states_upload(X)
        e_got, s_got = ctx.refine(e)
    print("e", np.abs(e_got - e_ref).max(), "sigma", np.abs(s_got - s_ref).max(), flush=True)
    assert np.abs(e_got - e_ref).max() < 1e-9
    assert np.abs(s_got - s_ref).max() < 1e-7


def test_ortho_fp32_ill_conditioned(sep):
    """sep_ortho_fp32 on 64 states of which 6 are nearly dependent: all 64
    kept, the basis orthonormal to 1e-12 and spanning the input, the
    singular values those of numpy's SVD to the
    FP64 Gram's resolution: eigenvalue errors of about eps sqrt(N) S_0^2
    (N = 32768 terms per entry), i.e. 1e-14 S_0^2 / S_i in S_i (7 % of the
    smallest here, whose ratios to S_0 are 4e-7 .. 1.3e-6)."""
    n, m = 32, 64
    rng = np.random.default_rng(7)
    A = rng.standard_normal((m, n ** 3))
    for r, eps in zip(range(58, 64), (1e-6, 1.5e-6, 2e-6, 2.5e-6, 3e-6, 1e-6)):
        A[r] = A[r - 58] + 0.5 * A[r - 57] + eps * rng.standard_normal(n ** 3)
    S_ref = np.linalg.svd(A, compute_uv=False)
    print("case ortho_fp32 ill-conditioned", S_ref[-1] / S_ref[0], flush=True)
    ctx = sep.Context(n, n, n, 0.5, 0.5, 0.5)
    try:
        ctx.states_upload(A)
        k, S = ctx.ortho_fp32(m)
        Q = ctx.states_download(k)
        dv = 0.125
    finally:
        ctx.close()
    assert k == m
    assert np.abs(Q @ Q.T * dv - np.eye(k)).max() < 1e-12
    # the span: A's rows are reproduced by their projection on Q
    P = (A @ Q.T) * dv
    assert np.abs(P @ Q - A).max() < 1e-8 * np.abs(A).max()
    tol = 1e-14 * S_ref[0] ** 2 / S_ref
    print("S max rel-to-tol", (np.abs(S - S_ref) / tol).max(), flush=True)
    assert np.all(np.abs(S - S_ref) <= tol + 1e-13 * S_ref[0])
