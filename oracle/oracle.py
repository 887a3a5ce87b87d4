"""ctypes wrapper over liboracle.so, the CPU restatement of the reference
`filter` program (filter/*.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_lp = ctypes.POINTER(ctypes.c_long)
_L = ctypes.c_long
_D = ctypes.c_double


def build():
    """Compile liboracle.so in place (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.or_fft3d.argtypes = [_L, _L, _L, _dp, ctypes.c_int]
        L.or_fft1d.argtypes = [_L, _dp, ctypes.c_int]
        L.or_ksqr.argtypes = [_L, _L, _L, _D, _D, _D, _D, _dp]
        L.or_build_potential.argtypes = [_L, _L, _L, _D, _D, _D, _D, _D, _D, _L,
                                         _dp, _dp, _dp, _lp, _dp, _lp, _dp, _dp,
                                         _L, _dp, _D, _D, _D, _dp]
        L.or_hf_adjoint.argtypes = [_L, _L, _L, _D, _D, _D, _D, _D, _D, _L, _dp, _dp, _dp,
                                    _lp, _dp, _lp, _dp, _dp, _L, _L, _dp, _dp, _D, _dp,
                                    _dp, _dp]
        L.or_interpolate.argtypes = [_D, _D, _dp, _dp, _L, _L, _L, _D]
        L.or_interpolate.restype = _D
        L.or_hamiltonian.argtypes = [_L, _L, _L, _dp, _dp, _dp, _L]
        L.or_ran_nrc.argtypes = [ctypes.c_void_p, _lp]
        L.or_ran_nrc.restype = _D
        L.or_rng_reset.argtypes = [ctypes.c_void_p]
        L.or_init_psi.argtypes = [ctypes.c_void_p, _lp, _L, _D, _dp]
        L.or_energy_range.argtypes = [_L, _L, _L, _dp, _dp, _D, _D, _dp, _dp, _lp]
        L.or_leja_points.argtypes = [_L, _dp]
        L.or_leja_points.restype = _D
        L.or_newton_coeffs.argtypes = [_L, _L, _dp, _dp, _D, _D, _D, _dp]
        L.or_filter_state.argtypes = [_L, _L, _L, _dp, _dp, _dp, _dp, _L, _L, _D,
                                      _D, _dp, _dp]
        L.or_set_interp.argtypes = [ctypes.c_int]
        L.or_filter_states.argtypes = [_L, _L, _L, _dp, _dp, _dp, _dp, _L, _L, _D, _D, _L, _dp,
                                       _dp]
        L.or_normalize_all.argtypes = [_dp, _D, _L, _L]
        L.or_energy_all.argtypes = [_L, _L, _L, _dp, _dp, _D, _dp, _L, _dp]
        L.or_ortho.argtypes = [_dp, _L, _L, _dp]
        L.or_ortho.restype = _L
        L.or_hmat.argtypes = [_L, _L, _L, _dp, _dp, _D, _dp, _L, _dp]
        L.or_sigma.argtypes = [_L, _L, _L, _dp, _dp, _D, _dp, _L, _dp]
        L.or_jacobi_eig.argtypes = [_L, _dp, _dp, _dp]
        L.or_run.argtypes = [ctypes.c_char_p, ctypes.c_char_p, _L, _L]
        L.or_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


def fft3d(vol, sign):
    """vol: complex128 array shaped (nz, ny, nx); returns a new array."""
    v = _c(vol, np.complex128).copy()
    nz, ny, nx = v.shape
    lib().or_fft3d(nx, ny, nz, v.ctypes.data_as(_dp), sign)
    return v


def fft1d(x, sign):
    v = _c(x, np.complex128).copy()
    lib().or_fft1d(v.size, v.ctypes.data_as(_dp), sign)
    return v


def ksqr(nx, ny, nz, dx, dy, dz, ekinmax=10.0):
    out = np.zeros(nx * ny * nz)
    lib().or_ksqr(nx, ny, nz, dx, dy, dz, ekinmax, _p(out))
    return out


def hamiltonian(grid, ksq, potl, phi):
    """phi: complex128 (batch, N) or (N,); returns H phi (new array)."""
    nx, ny, nz = grid
    v = _c(phi, np.complex128).copy()
    batch = 1 if v.ndim == 1 else v.shape[0]
    lib().or_hamiltonian(nx, ny, nz, _p(_c(ksq)), _p(_c(potl)), v.ctypes.data_as(_dp), batch)
    return v


class Rng(ctypes.Structure):
    _fields_ = [("idum2", ctypes.c_long), ("iy", ctypes.c_long), ("iv", ctypes.c_long * 32)]

    def __init__(self):
        super().__init__()
        lib().or_rng_reset(ctypes.byref(self))

    def ran(self, idum):
        """Draw one number; idum is a 1-element np.int64 array (updated)."""
        return lib().or_ran_nrc(ctypes.byref(self), idum.ctypes.data_as(_lp))

    def init_psi(self, idum, ngrid, dv):
        out = np.zeros(ngrid)
        lib().or_init_psi(ctypes.byref(self), idum.ctypes.data_as(_lp), ngrid, dv, _p(out))
        return out


def energy_range(grid, ksq, potl, dv, vmin):
    nx, ny, nz = grid
    dE = np.zeros(1)
    emax = np.zeros(1)
    it = np.zeros(1, dtype=np.int64)
    lib().or_energy_range(nx, ny, nz, _p(_c(ksq)), _p(_c(potl)), dv, vmin, _p(dE),
                          _p(emax), it.ctypes.data_as(_lp))
    return float(dE[0]), float(emax[0]), int(it[0])


def leja_points(nc):
    zn = np.zeros(nc)
    fk = lib().or_leja_points(nc, _p(zn))
    return zn, fk


def newton_coeffs(nc, el, zn, dE, vmin, dt):
    el = _c(el)
    an = np.zeros((el.size, nc))
    lib().or_newton_coeffs(nc, el.size, _p(_c(zn)), _p(el), dE, vmin, dt, _p(an))
    return an


def filter_state(grid, ksq, potl, an, zn, dE_1, vmin, psi0):
    nx, ny, nz = grid
    an = _c(an)
    ms, nc = an.shape
    out = np.zeros((ms, nx * ny * nz))
    lib().or_filter_state(nx, ny, nz, _p(_c(ksq)), _p(_c(potl)), _p(an), _p(_c(zn)), nc,
                          ms, dE_1, vmin, _p(_c(psi0)), _p(out))
    return out


def set_interp(interp):
    """0 = linear table interpolation (the reference), 1 = cubic Hermite."""
    lib().or_set_interp(int(interp))


def filter_states(grid, ksq, potl, an, zn, dE_1, vmin, psi0):
    """or_filter_state for every row of psi0 (ns x N), states in parallel;
    returns (ns, ms, N)."""
    nx, ny, nz = grid
    an = _c(an)
    ms, nc = an.shape
    p0 = _c(psi0).reshape(-1, nx * ny * nz)
    out = np.zeros((p0.shape[0], ms, nx * ny * nz))
    lib().or_filter_states(nx, ny, nz, _p(_c(ksq)), _p(_c(potl)), _p(an), _p(_c(zn)), nc, ms,
                           dE_1, vmin, p0.shape[0], _p(p0), _p(out))
    return out


def normalize_all(psi, dv):
    p = _c(psi).copy()
    m = p.shape[0] if p.ndim > 1 else 1
    lib().or_normalize_all(_p(p), dv, m, p.size // m)
    return p


def energy_all(grid, ksq, potl, dv, psi):
    nx, ny, nz = grid
    p = _c(psi)
    ene = np.zeros(p.shape[0])
    lib().or_energy_all(nx, ny, nz, _p(_c(ksq)), _p(_c(potl)), dv, _p(p), p.shape[0], _p(ene))
    return ene


def ortho(psi):
    """psi: (m, N) states; returns (U (m, N), S (m,), cutoff)."""
    p = _c(psi).copy()
    m, n = p.shape
    S = np.zeros(m)
    cut = lib().or_ortho(_p(p), n, m, _p(S))
    return p, S, int(cut)


def hmat(grid, ksq, potl, dv, psi):
    nx, ny, nz = grid
    p = _c(psi).copy()
    ev = np.zeros(p.shape[0])
    lib().or_hmat(nx, ny, nz, _p(_c(ksq)), _p(_c(potl)), dv, _p(p), p.shape[0], _p(ev))
    return ev, p


def sigma(grid, ksq, potl, dv, psi):
    nx, ny, nz = grid
    p = _c(psi)
    s = np.zeros(p.shape[0])
    lib().or_sigma(nx, ny, nz, _p(_c(ksq)), _p(_c(potl)), dv, _p(p), p.shape[0], _p(s))
    return s


def jacobi_eig(A):
    A = _c(A)
    n = A.shape[0]
    w = np.zeros(n)
    V = np.zeros((n, n))
    lib().or_jacobi_eig(n, _p(A), _p(w), _p(V))
    return w, V


def build_potential(grid, origin, spacing, atoms_xyz, types, scale, npot, vr, pot, dr,
                    expot=(0.0, 0.0, 1.0)):
    nx, ny, nz = grid
    xyz = _c(atoms_xyz)
    rx, ry, rz = _c(xyz[:, 0]), _c(xyz[:, 1]), _c(xyz[:, 2])
    t = _c(types, np.int64)
    sc = _c(scale)
    npot = _c(npot, np.int64)
    vr = _c(vr)
    pot = _c(pot)
    out = np.zeros(nx * ny * nz)
    lib().or_build_potential(nx, ny, nz, origin[0], origin[1], origin[2], spacing[0],
                             spacing[1], spacing[2], xyz.shape[0], _p(rx), _p(ry), _p(rz),
                             t.ctypes.data_as(_lp), _p(sc), npot.ctypes.data_as(_lp), _p(vr),
                             _p(pot), vr.shape[1], _p(_c(dr)), expot[0], expot[1], expot[2],
                             _p(out))
    return out


def hf_adjoint(grid, origin, spacing, atoms_xyz, types, scale, npot, vr, pot, dr, rho, dv):
    """(dR natom x 3, dscale natom, dtab ntype x stride) of one density rho."""
    nx, ny, nz = grid
    xyz = _c(atoms_xyz)
    rx, ry, rz = _c(xyz[:, 0]), _c(xyz[:, 1]), _c(xyz[:, 2])
    t = _c(types, np.int64)
    sc = _c(scale)
    npot = _c(npot, np.int64)
    vr = _c(vr)
    pot = _c(pot)
    natom = xyz.shape[0]
    dR = np.zeros((natom, 3))
    ds = np.zeros(natom)
    dtab = np.zeros(vr.shape)
    lib().or_hf_adjoint(nx, ny, nz, origin[0], origin[1], origin[2], spacing[0], spacing[1],
                        spacing[2], natom, _p(rx), _p(ry), _p(rz), t.ctypes.data_as(_lp),
                        _p(sc), npot.ctypes.data_as(_lp), _p(vr), _p(pot), vr.shape[1],
                        vr.shape[0], _p(_c(dr)), _p(_c(rho)), dv, _p(dR), _p(ds), _p(dtab))
    return dR, ds, dtab


def run(indir, outdir, idum, nthreads=0):
    """Whole reference pipeline on a deck; returns the summary dict."""
    os.makedirs(outdir, exist_ok=True)
    rc = lib().or_run(indir.encode(), outdir.encode(), idum, nthreads)
    if rc != 0:
        raise RuntimeError(lib().or_last_error().decode())
    summ = {}
    with open(os.path.join(outdir, "oracle_summary.dat")) as f:
        for line in f:
            k, v = line.split()
            summ[k] = float(v)
    return summ
