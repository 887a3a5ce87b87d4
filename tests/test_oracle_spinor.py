"""The spinor oracle (oracle/spinor.py) pinned by what the mathematics fixes:
its nonlocal spin-orbit operator against an independently assembled dense
matrix of the reference's block form (spinors/filterSpinors/hamiltonian.c:
160-185), Hermiticity, reduction to the scalar H (pinned against the
reference's In13_P16 deck elsewhere) when Vso = 0, the reference's sign of
the down-channel m = -1 term (which breaks time-reversal symmetry), and the
whole filter -> ortho -> Rayleigh-Ritz pipeline against eigenvalues of the
dense spinor H on a 10^3 grid."""
import numpy as np
import pytest

N1 = 10
D = 0.6


@pytest.fixture(scope="module")
def tiny(oracle):
    from oracle import spinor as S
    grid = (N1, N1, N1)
    origin = (-N1 * D / 2,) * 3
    xyz = np.array([[0.1, -0.2, 0.15], [1.2, 0.9, -0.7]])
    vso = np.array([0.8, 0.5])
    lists = S.so_lists(grid, origin, (D, D, D), xyz, vso)
    ksq = oracle.ksqr(N1, N1, N1, D, D, D)
    V = np.random.default_rng(1).uniform(-1.0, 0.3, N1 ** 3)
    return S, grid, origin, xyz, vso, lists, ksq, V


def _rand(rng, N):
    return rng.standard_normal((2, N)) + 1j * rng.standard_normal((2, N))


def test_lists_follow_the_cutoff_sphere(tiny):
    S, grid, origin, xyz, vso, lists, ksq, V = tiny
    ax = S.grid_axis(origin[0], D, N1)
    for a, e in enumerate(lists):
        kz, r = np.divmod(e["jxyz"], N1 * N1)
        ky, kx = np.divmod(r, N1)
        d2 = (ax[kx] - xyz[a, 0]) ** 2 + (ax[ky] - xyz[a, 1]) ** 2 + (ax[kz] - xyz[a, 2]) ** 2
        assert np.all(d2 < S.RNLCUT2) and np.all(np.diff(e["jxyz"]) > 0)
        # every grid point inside the sphere is listed (brute force)
        Z, Y, X = np.meshgrid(ax, ax, ax, indexing="ij")
        inside = ((X - xyz[a, 0]) ** 2 + (Y - xyz[a, 1]) ** 2 + (Z - xyz[a, 2]) ** 2) < S.RNLCUT2
        assert inside.sum() == e["jxyz"].size
        np.testing.assert_allclose(e["vr"], vso[a] * np.exp(-d2 / 0.49), rtol=1e-14)
        # |Y_1m|^2 summed over m is 3 / (4 pi) (up to the EPSDX shift)
        s = np.abs(e["y10"]) ** 2 + np.abs(e["y11"]) ** 2 + np.abs(e["y1_1"]) ** 2
        np.testing.assert_allclose(s[e["r"] > 1e-3], 3.0 / (4.0 * np.pi), rtol=1e-12)


def test_so_apply_equals_dense_block_form(tiny):
    S, grid, origin, xyz, vso, lists, ksq, V = tiny
    N = N1 ** 3
    M = S.so_dense(lists, N, D ** 3, D)
    a = _rand(np.random.default_rng(2), N)
    got = S.so_apply(lists, a, D ** 3, D).reshape(-1)
    assert np.abs(got - M @ a.reshape(-1)).max() < 1e-14 * np.abs(got).max()
    assert np.abs(M - M.conj().T).max() < 1e-15 * np.abs(M).max()


def test_spinor_h_hermitian_and_scalar_limit(tiny, oracle):
    S, grid, origin, xyz, vso, lists, ksq, V = tiny
    N = N1 ** 3
    rng = np.random.default_rng(3)
    a, b = _rand(rng, N), _rand(rng, N)
    Ha = S.hamiltonian(grid, ksq, V, lists, a, D ** 3, D)
    Hb = S.hamiltonian(grid, ksq, V, lists, b, D ** 3, D)
    assert abs(np.vdot(a, Hb) - np.conj(np.vdot(b, Ha))) < 1e-13 * abs(np.vdot(a, Hb))
    zero = S.so_lists(grid, origin, (D, D, D), xyz, np.zeros(len(xyz)))
    H0 = S.hamiltonian(grid, ksq, V, zero, a, D ** 3, D)
    ref = oracle.hamiltonian(grid, ksq, V, a)
    assert np.array_equal(H0, ref)


def test_reference_sign_breaks_time_reversal(tiny):
    """With the down-channel term -t1_1 S1_1dn of hamiltonian.c:179-181 the
    operator does not commute with Theta psi = (-conj dn, conj up); flipping
    that sign (the L.S form) would make it commute.  Pins the sign."""
    S, grid, origin, xyz, vso, lists, ksq, V = tiny
    N = N1 ** 3
    a = _rand(np.random.default_rng(4), N)
    th = lambda p: np.array([-np.conj(p[1]), np.conj(p[0])])
    so = lambda p: S.so_apply(lists, p, D ** 3, D)
    rel = np.abs(so(th(a)) - th(so(a))).max() / np.abs(so(a)).max()
    assert rel > 1e-2
    # the same operator with that one term's sign flipped commutes with Theta
    c = 0.5 / np.sqrt(2.0 * np.pi * D * D) * D ** 3
    fix = np.zeros_like(a)
    for e in lists:
        K = S.so_kernel(e, D)
        j = e["jxyz"]
        s = K @ (np.conj(e["y1_1"]) * a[1, j])
        fix[1, j] += 2.0 * c * e["y1_1"] * s
    lf = lambda p, f: S.so_apply(lists, p, D ** 3, D) + f
    fix_t = np.zeros_like(a)
    ta = th(a)
    for e in lists:
        K = S.so_kernel(e, D)
        j = e["jxyz"]
        fix_t[1, j] += 2.0 * c * e["y1_1"] * (K @ (np.conj(e["y1_1"]) * ta[1, j]))
    rel_fixed = np.abs(lf(ta, fix_t) - th(lf(a, fix))).max() / np.abs(so(a)).max()
    assert rel_fixed < 1e-13


def test_pipeline_matches_dense_eigenvalues(tiny, oracle):
    """Energy range, filter, portho and Hmat on the 10^3 grid: the converged
    Ritz values (sigma < 1e-6) equal eigenvalues of the dense spinor H to
    1e-9 Ha."""
    S, grid, origin, xyz, vso, lists, ksq, V = tiny
    N = N1 ** 3
    dv = D ** 3
    M = S.so_dense(lists, N, dv, D)
    TV = oracle.hamiltonian(grid, ksq, V, np.eye(N, dtype=np.complex128))
    Hd = M.copy()
    Hd[:N, :N] += TV.T
    Hd[N:, N:] += TV.T
    w = np.linalg.eigvalsh(Hd)
    vmin, dE, its = S.energy_range(grid, ksq, V, lists, dv, D)
    assert vmin < w[0] and vmin + dE > w[-1]
    nc, ms, ns = 512, 4, 8
    zn, _ = oracle.leja_points(nc)
    el = np.linspace(w[0] - 0.05, w[5] + 0.02, ms)
    an = oracle.newton_coeffs(nc, el, zn, dE, vmin, (nc / (2.5 * dE)) ** 2)
    rng = oracle.Rng()
    idum = np.array([-1001], dtype=np.int64)
    filt = np.concatenate([S.filter_spinor(grid, ksq, V, lists, dv, D, an, zn, dE, vmin,
                                           S.init_spinor(rng, idum, N, dv)) for _ in range(ns)])
    U, sv, k = S.portho(filt)
    U = np.array([u / np.sqrt(S.spinor_norm2(u, dv)) for u in U])
    ev, psi = S.hmat(grid, ksq, V, lists, dv, D, U)
    sg = S.sigma_e(grid, ksq, V, lists, dv, D, psi)
    conv = sg < 1e-6
    err = np.array([np.abs(w - e).min() for e in ev[conv]])
    print("spinor pipeline: kept", k, "converged", conv.sum(), "max |dE|", err.max(),
          "sigma", np.array2string(sg, precision=1))
    assert conv.sum() >= 4
    assert err.max() < 1e-9, err
