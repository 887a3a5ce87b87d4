"""Oracle subspace stage: ortho (filter/ortho.c:11-52, MKL dgesvd), Rayleigh-Ritz
(filter/Hmat.c:11-99, MKL dsyev) and sigma (filter/energy.c:88-119), pinned
by numpy.linalg / scipy and closed forms."""
import numpy as np
import scipy.linalg as sla


def test_jacobi_eig_vs_numpy(oracle):
    rng = np.random.default_rng(1)
    A = rng.standard_normal((40, 40))
    A = A + A.T
    w, V = oracle.jacobi_eig(A)
    wr = np.linalg.eigvalsh(A)
    assert np.abs(w - wr).max() < 1e-12 * np.abs(wr).max()
    assert np.abs(A @ V - V * w).max() < 1e-11 * np.abs(wr).max()


def test_ortho_vs_numpy_svd(oracle):
    rng = np.random.default_rng(2)
    m, n = 12, 500
    # singular values spanning 1 .. 1e-12, plus one exactly dependent state
    U0, _ = np.linalg.qr(rng.standard_normal((n, m)))
    V0, _ = np.linalg.qr(rng.standard_normal((m, m)))
    s0 = np.array([1, .5, .1, 1e-2, 1e-4, 1e-6, 1e-8, 1e-9, 3e-10, 5e-11, 1e-12, 0.0])
    A = (U0 * s0) @ V0.T  # n x m, columns are states
    U, S, cut = oracle.ortho(A.T)
    Sr = np.linalg.svd(A, compute_uv=False)
    assert np.abs(S - Sr).max() < 1e-14 * Sr[0]
    assert cut == 9  # first S_i/S_0 < 1e-10 (ortho.c:41-46): 5e-11
    # kept columns orthonormal and spanning the leading left singular space
    K = U[:6]
    assert np.abs(K @ K.T - np.eye(6)).max() < 1e-12
    P = U0[:, :6]
    assert np.abs(K.T - P @ (P.T @ K.T)).max() < 1e-9


def test_hmat_vs_scipy(oracle):
    grid, d = (10, 10, 10), (0.8, 0.8, 0.8)
    rng = np.random.default_rng(3)
    N = 1000
    pot = rng.uniform(-1, 0.5, N)
    ksq = oracle.ksqr(*grid, *d)
    dv = np.prod(d)
    X, _ = np.linalg.qr(rng.standard_normal((N, 8)))
    psi = X.T / np.sqrt(dv)  # normalised with dv
    ev, rot = oracle.hmat(grid, ksq, pot, dv, psi)
    Hpsi = oracle.hamiltonian(grid, ksq, pot, psi.astype(complex)).real
    Hm = (psi @ Hpsi.T) * dv
    ref = sla.eigh(0.5 * (Hm + Hm.T), eigvals_only=True)
    assert np.abs(ev - ref).max() < 1e-11
    # rotated states are Ritz vectors: residual orthogonal to the subspace
    R = oracle.hamiltonian(grid, ksq, pot, rot.astype(complex)).real - ev[:, None] * rot
    assert np.abs((rot @ R.T) * dv).max() < 1e-10


def test_sigma_zero_for_eigenvectors(oracle):
    """Plane waves with constant V are exact eigenvectors: sigma = 0."""
    grid, d = (8, 8, 8), (0.7, 0.7, 0.7)
    N = 512
    ksq = oracle.ksqr(*grid, *d)
    pot = np.full(N, -0.2)
    x = np.arange(8) * 0.7
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    k = 2 * np.pi / (8 * 0.7)
    cosw = np.cos(k * X).reshape(-1)
    dv = 0.7 ** 3
    cosw /= np.sqrt(np.sum(cosw ** 2) * dv)
    s = oracle.sigma(grid, ksq, pot, dv, cosw[None, :])
    assert s[0] < 1e-6
    mix = cosw + np.cos(2 * k * Y).reshape(-1) / np.sqrt(np.sum(np.cos(2 * k * Y) ** 2) * dv)
    mix /= np.sqrt(np.sum(mix ** 2) * dv)
    e1, e2 = 0.5 * k * k - 0.2, 0.5 * (2 * k) ** 2 - 0.2
    s2 = oracle.sigma(grid, ksq, pot, dv, mix[None, :])
    assert abs(s2[0] - 0.5 * abs(e2 - e1)) < 1e-10
