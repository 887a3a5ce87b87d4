"""CPU restatement of the reference's spin-orbit spinor filter
(spinors/filterSpinors/*.c) with the matrix form of the nonlocal spin-orbit
potential (flagkb = 0, hamiltonian.c:79-192).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker.  The product
never imports it.

A spinor is a complex128 array (2, N): the up component, then the down
component (the reference's zomplex[2 ngrid], hamiltonian.c:18-19).  The
kinetic and local parts act on each component as the scalar H of
oracle.hamiltonian (filter/hamiltonian.c), which is pinned separately.

Parity pinning (tests/test_oracle_spinor.py): the operator is Hermitian
(<a|H b> = conj <b|H a>), it reduces to the scalar H on each component when
every Vso is zero, its nonlocal part equals a dense matrix assembled
independently from the block form sum_pq |p> K(p,q) Y_m(p) C^{ss'}_{mm'}
conj(Y_m'(q)) <q| of the reference's formula, and the spinor pipeline
(filter -> ortho -> Rayleigh-Ritz) on a tiny grid reproduces the lowest
eigenvalues of the dense spinor H (numpy.linalg.eigh).  The reference holds
no spinor fixtures, so nothing pins it against a run of the reference:
parity unpinned beyond these checks.
"""
import functools

import numpy as np
from threadpoolctl import threadpool_limits

from . import oracle as O


def _serial(f):
    """Runs f with one BLAS / OpenMP thread: its products are small, and the
    libraries' thread pools spinning against each other cost 25x."""
    @functools.wraps(f)
    def g(*a, **k):
        with threadpool_limits(limits=1):
            return f(*a, **k)
    return g

RNLCUT2 = 0.49 * 6.0 * np.log(10.0)  # init.c:99
EPSDX = 1.0e-20  # fd.h:100
SQRT2 = np.sqrt(2.0)


def grid_axis(xmin, d, n):
    """vx[j] as the reference accumulates it (init.c:467-469)."""
    v = np.empty(n)
    x = xmin
    for j in range(n):
        v[j] = x
        x += d
    return v


def so_lists(grid, origin, spacing, xyz, vso):
    """init_list (spinors/filterSpinors/init.c:114-183): per atom the grid
    points with |r - R|^2 < Rnlcut2, in z, y, x order, with Vr, r, r2 and the
    l = 1 spherical harmonics y10, y11, y1_1 (EPSDX-shifted)."""
    nx, ny, nz = grid
    vx = grid_axis(origin[0], spacing[0], nx)
    vy = grid_axis(origin[1], spacing[1], ny)
    vz = grid_axis(origin[2], spacing[2], nz)
    tmp1 = 0.5 * np.sqrt(3.0 / np.pi)
    tmp2 = 0.5 * np.sqrt(3.0 / (2.0 * np.pi))
    tmp49 = 1.0 / 0.49
    lists = []
    for a in range(len(xyz)):
        dz = vz[:, None, None] - xyz[a, 2]
        dy = vy[None, :, None] - xyz[a, 1]
        dx = vx[None, None, :] - xyz[a, 0]
        dr2 = dx * dx + dy * dy + dz * dz
        mask = dr2 < RNLCUT2
        kz, ky, kx = np.nonzero(mask)  # C order: z outer, x inner
        ddx = vx[kx] - xyz[a, 0]
        ddy = vy[ky] - xyz[a, 1]
        ddz = vz[kz] - xyz[a, 2]
        d2 = ddx * ddx + ddy * ddy + ddz * ddz
        dxe, dye, dze = ddx + EPSDX, ddy + EPSDX, ddz + EPSDX
        e = {}
        e["jxyz"] = (kz * ny + ky) * nx + kx
        e["vr"] = vso[a] * np.exp(-tmp49 * d2)
        e["r"] = np.sqrt(d2)
        y10 = tmp1 * dze / np.sqrt(ddx * ddx + ddy * ddy + dze * dze)
        rx_1 = 1.0 / np.sqrt(dxe * dxe + ddy * ddy + ddz * ddz)
        ry_1 = 1.0 / np.sqrt(ddx * ddx + dye * dye + ddz * ddz)
        e["y10"] = y10 + 0j
        e["y11"] = -tmp2 * dxe * rx_1 - 1j * tmp2 * dye * ry_1
        e["y1_1"] = tmp2 * dxe * rx_1 - 1j * tmp2 * dye * ry_1
        e["r2"] = np.where(d2 > EPSDX, d2, 1.0 / EPSDX)
        lists.append(e)
    return lists


def so_kernel(e, sigma):
    """K(p,q) = <Vso>_pq exp(-((r_p - r_q) sigma_1)^2) / <r^2>_pq
    (hamiltonian.c:123-130), sigma_1 = sqrt(1/2) / sigma (init.c:97);
    cached in the list entry per sigma."""
    key = ("K", sigma)
    if key not in e:
        s1 = np.sqrt(0.5) / sigma
        rr = e["r"][:, None] - e["r"][None, :]
        e[key] = (0.5 * (e["vr"][:, None] + e["vr"][None, :]) * np.exp(-(rr * s1) ** 2) /
                  (0.5 * (e["r2"][:, None] + e["r2"][None, :])))
    return e[key]


def so_apply(lists, psi, dv, sigma):
    """nonlocal_potential (hamiltonian.c:79-192): V_so psi for one spinor."""
    up, dn = psi
    out = np.zeros_like(psi)
    c = 0.5 / np.sqrt(2.0 * np.pi * sigma * sigma) * dv
    for e in lists:
        if e["jxyz"].size == 0:
            continue
        K = so_kernel(e, sigma)
        j = e["jxyz"]
        S = {}
        for m in ("y11", "y10", "y1_1"):
            yc = np.conj(e[m])
            S[m, "up"] = K @ (yc * up[j])
            S[m, "dn"] = K @ (yc * dn[j])
        t11, t10, t1_1 = c * e["y11"], c * e["y10"], c * e["y1_1"]
        out[0, j] += (t11 * S["y11", "up"] - t1_1 * S["y1_1", "up"] +
                      SQRT2 * (t10 * S["y11", "dn"] + t1_1 * S["y10", "dn"]))
        out[1, j] += (-t11 * S["y11", "dn"] - t1_1 * S["y1_1", "dn"] +
                      SQRT2 * (t10 * S["y1_1", "up"] + t11 * S["y10", "up"]))
    return out


@_serial
def so_dense(lists, N, dv, sigma):
    """The nonlocal operator as a dense (2N, 2N) matrix from its block form:
    sum over atoms and p, q of |p,s> c K(p,q) Y_m(p) C[s,m ; s',m'] conj(Y_m'(q)) <q,s'|,
    C read off hamiltonian.c:160-185.  For small grids (checks only)."""
    M = np.zeros((2 * N, 2 * N), dtype=np.complex128)
    c = 0.5 / np.sqrt(2.0 * np.pi * sigma * sigma) * dv
    # (s, m, s', m', coefficient): up = 0, down = 1
    terms = [(0, "y11", 0, "y11", 1.0), (0, "y1_1", 0, "y1_1", -1.0),
             (0, "y10", 1, "y11", SQRT2), (0, "y1_1", 1, "y10", SQRT2),
             (1, "y11", 1, "y11", -1.0), (1, "y1_1", 1, "y1_1", -1.0),
             (1, "y10", 0, "y1_1", SQRT2), (1, "y11", 0, "y10", SQRT2)]
    for e in lists:
        if e["jxyz"].size == 0:
            continue
        K = so_kernel(e, sigma)
        j = e["jxyz"]
        for s, m, s2, m2, coef in terms:
            blk = c * coef * (e[m][:, None] * K * np.conj(e[m2])[None, :])
            M[np.ix_(s * N + j, s2 * N + j)] += blk
    return M


def hamiltonian(grid, ksq, potl, lists, psi, dv, sigma):
    """hamiltonian (spinors/filterSpinors/hamiltonian.c:8-50) on one spinor (2, N)."""
    tv = O.hamiltonian(grid, ksq, potl, np.asarray(psi, dtype=np.complex128))
    return tv + so_apply(lists, psi, dv, sigma)


def init_spinor(rng, idum, N, dv):
    """One random spinor: init_psi on the up, then the down component
    (main.c:87-89, init.c:546-563: real parts random, imaginary parts zero,
    each component normalized)."""
    return np.array([rng.init_psi(idum, N, dv), rng.init_psi(idum, N, dv)], dtype=np.complex128)


def spinor_norm2(psi, dv):
    return float(np.sum(np.abs(psi) ** 2) * dv)


@_serial
def energy_range(grid, ksq, potl, lists, dv, sigma, tau=0.025):
    """get_energy_range (spinors/filterSpinors/energy.c:53-112): Emin by
    phi <- (1 - tau H) phi, Emax by phi <- H phi, each normalized, until the
    relative change of <phi|H phi> dv is below 1e-6 (at most 500); then
    Emax *= 1.2, Emin -= 0.2 |Emin|.  Returns (Vmin = Emin, dE, iters)."""
    N = int(np.prod(grid))
    rng = O.Rng()
    idum = np.array([-874917403], dtype=np.int64)
    H = lambda p: hamiltonian(grid, ksq, potl, lists, p, dv, sigma)
    # H of the normalized iterate is that of the next step: one H per step
    phi = init_spinor(rng, idum, N, dv)
    hphi = H(phi)
    emin, old, i1 = 0.1, 0.0, 0
    while abs((emin - old) / emin) > 1.0e-6 and i1 < 500:
        phi = phi - tau * hphi
        phi /= np.sqrt(spinor_norm2(phi, dv))
        hphi = H(phi)
        old = emin
        emin = float(np.sum(np.conj(phi) * hphi).real * dv)
        i1 += 1
    psi = init_spinor(rng, idum, N, dv)
    hpsi = H(psi)
    emax, old, i2 = 0.1, 0.0, 0
    while abs((emax - old) / emax) > 1.0e-6 and i2 < 500:
        psi = hpsi / np.sqrt(spinor_norm2(hpsi, dv))
        hpsi = H(psi)
        old = emax
        emax = float(np.sum(np.conj(psi) * hpsi).real * dv)
        i2 += 1
    emax *= 1.2
    emin -= 0.2 * abs(emin)
    return emin, emax - emin, (i1, i2)


@_serial
def filter_spinor(grid, ksq, potl, lists, dv, sigma, an, zn, dE, vmin, psi0):
    """filter (spinors/filterSpinors/filter.c:57-83) on one spinor: ms
    filtered spinors (ms, 2, N), then normalize_all (filter.c:30)."""
    an = np.asarray(an, dtype=np.float64)
    ms, nc = an.shape
    dE_1 = 4.0 / dE
    psin = np.array(psi0, dtype=np.complex128)
    out = an[:, 0][:, None, None] * psin[None]
    for jc in range(1, nc):
        psim1 = psin
        h = hamiltonian(grid, ksq, potl, lists, psim1, dv, sigma)
        psin = dE_1 * h - (2.0 + zn[jc - 1] + vmin * dE_1) * psim1
        out += an[:, jc][:, None, None] * psin[None]
    for e in range(ms):
        out[e] /= np.sqrt(spinor_norm2(out[e], dv))
    return out


@_serial
def portho(psi, cutoff=1.0e-10):
    """portho (spinors/filterSpinors/ortho.c:5-45): left singular vectors of
    the (2N x m) matrix of spinors, kept while S_i / S_0 >= cutoff; returns
    (U (k, 2, N) normalized as main.c:137, S, k)."""
    m = psi.shape[0]
    A = psi.reshape(m, -1).T
    U, S, _ = np.linalg.svd(A, full_matrices=False)
    k = m
    for i in range(m):
        if S[i] / S[0] < cutoff:
            k = i
            break
    return U[:, :k].T.reshape(k, 2, -1), S, k


@_serial
def hmat(grid, ksq, potl, lists, dv, sigma, psi):
    """Hmat (spinors/filterSpinors/Hmat.c:5-72): eigenvalues ascending and
    the Ritz spinors (normalized as main.c:146)."""
    k = psi.shape[0]
    hp = np.array([hamiltonian(grid, ksq, potl, lists, p, dv, sigma) for p in psi])
    A = psi.reshape(k, -1)
    P = np.conj(A) @ hp.reshape(k, -1).T * dv  # P_ij = <psi_i|H|psi_j>
    P = 0.5 * (P + np.conj(P.T))
    w, Z = np.linalg.eigh(P)
    out = (Z.T @ A).reshape(psi.shape)
    for i in range(k):
        out[i] /= np.sqrt(spinor_norm2(out[i], dv))
    return w, out


@_serial
def sigma_e(grid, ksq, potl, lists, dv, sigma, psi):
    """calc_sigma_E (spinors/filterSpinors/energy.c:117-147)."""
    out = np.empty(psi.shape[0])
    for i, p in enumerate(psi):
        h = hamiltonian(grid, ksq, potl, lists, p, dv, sigma)
        h2 = hamiltonian(grid, ksq, potl, lists, h, dv, sigma)
        e = float(np.sum(np.conj(p) * h).real * dv)
        e2 = float(np.sum(np.conj(p) * h2).real * dv)
        out[i] = np.sqrt(abs(e2 - e * e))
    return out
