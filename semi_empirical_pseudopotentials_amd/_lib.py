"""ctypes binding of libsep (include/sep.h).

The Python side is a thin host mirror of the C ABI; every numerical call runs
in libsep's HIP kernels.  There is no CPU fallback: importing works anywhere,
but constructing a Context without the built library or without a GPU raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libsep.so")
FILTER_BIN = os.path.join(_HERE, "bin", "filter")
COUPLINGS_BIN = os.path.join(_HERE, "bin", "couplings")
FILTER_SPINORS_BIN = os.path.join(_HERE, "bin", "filter_spinors")

SEP_OK = 0
SEP_INTERP_LINEAR = 0
SEP_INTERP_CUBIC = 1
SEP_HPSI_PLANE = 0
SEP_HPSI_PLANE_CHUNKED = 1
SEP_HPSI_LINES = 2
SEP_HPSI_PLANE_SWIZZLED = 3
SEP_HPSI_NFAMILY = 4
SEP_PREC_FP64 = 0
SEP_PREC_FP32 = 1
_ERRNAMES = {1: "SEP_ERR_ARG", 2: "SEP_ERR_UNSUPPORTED", 3: "SEP_ERR_HIP", 4: "SEP_ERR_IO",
             5: "SEP_ERR_NOMEM", 6: "SEP_ERR_NUMERIC"}

_dp = ctypes.POINTER(ctypes.c_double)
_lp = ctypes.POINTER(ctypes.c_long)
_ip = ctypes.POINTER(ctypes.c_int)
_L = ctypes.c_long
_D = ctypes.c_double
_I = ctypes.c_int
_V = ctypes.c_void_p


class SepError(RuntimeError):
    pass


class GridDesc(ctypes.Structure):
    _fields_ = [("nx", _L), ("ny", _L), ("nz", _L), ("dx", _D), ("dy", _D), ("dz", _D),
                ("xmin", _D), ("ymin", _D), ("zmin", _D), ("ekinmax", _D), ("device", _I)]


NTYPE = 17
NPOT = 8192


class Deck(ctypes.Structure):
    """sep_deck (include/sep_deck.h)."""
    _fields_ = [("nx", _L), ("ny", _L), ("nz", _L), ("ms", _L), ("ns", _L), ("nc", _L),
                ("nthreads", _L), ("Elmin", _D), ("Elmax", _D), ("natom", _L),
                ("xyz", _dp), ("type", _ip), ("sym", ctypes.c_void_p), ("scale", _dp),
                ("has_strain", _I), ("expot", _D * 3), ("has_expot", _I),
                ("npot", _L * NTYPE), ("dr", _D * NTYPE), ("vr", _dp), ("pot", _dp),
                ("xd", _D), ("yd", _D), ("zd", _D), ("xmin", _D), ("ymin", _D), ("zmin", _D),
                ("dx", _D), ("dy", _D), ("dz", _D), ("dv", _D), ("seed", _L), ("has_seed", _I)]


class Rng(ctypes.Structure):
    _fields_ = [("idum2", _L), ("iy", _L), ("iv", _L * 32)]


# Exported symbols and their signatures (restype, argtypes); the loader
# checks every one of them, the tests check the list against include/*.h.
SIGNATURES = {
    "sep_open": (_I, [ctypes.POINTER(GridDesc), ctypes.POINTER(_V)]),
    "sep_close": (_I, [_V]),
    "sep_last_error": (ctypes.c_char_p, [_V]),
    "sep_fft_length_supported": (_I, [_L]),
    "sep_set_potential": (_I, [_V, _dp]),
    "sep_build_potential": (_I, [_V, _L, _dp, _ip, _dp, _I, _lp, _dp, _dp, _L, _dp, _dp, _dp,
                                 _dp, _dp]),
    "sep_apply_h": (_I, [_V, _dp, _dp, _L]),
    "sep_hnorm": (_I, [_V, _dp, _dp, _L, _D, _D, _D]),
    "sep_apply_h_dev": (_I, [_V, _V, _V, _L, _I]),
    "sep_apply_h_dev_noalloc": (_I, [_V, _V, _V, _L, _I]),
    "sep_set_interp": (_I, [_V, _I]),
    "sep_energy_range": (_I, [_V, _D, _dp, _dp, _lp]),
    "sep_rng_reset": (None, [ctypes.POINTER(Rng)]),
    "sep_ran_nrc": (_D, [ctypes.POINTER(Rng), _lp]),
    "sep_init_psi": (None, [ctypes.POINTER(Rng), _lp, _L, _D, _dp]),
    "sep_leja_points": (_D, [_L, _dp]),
    "sep_newton_coeffs": (_I, [_L, _L, _dp, _dp, _D, _D, _D, _dp]),
    "sep_filter": (_I, [_V, _dp, _L, _dp, _dp, _L, _L, _D, _D, _I, _dp, _dp]),
    "sep_states_upload": (_I, [_V, _dp, _L]),
    "sep_states_download": (_I, [_V, _dp, _L]),
    "sep_normalize_all": (_I, [_V, _L]),
    "sep_energy_all": (_I, [_V, _L, _dp]),
    "sep_ortho": (_I, [_V, _L, _D, _dp, _lp]),
    "sep_rayleigh_ritz": (_I, [_V, _L, _dp]),
    "sep_sigma": (_I, [_V, _L, _dp]),
    "sep_ortho_fp32": (_I, [_V, _L, _D, _dp, _lp]),
    "sep_hf_adjoint": (_I, [_V, _L, _lp, _dp, _L, _dp, ctypes.POINTER(ctypes.c_int), _dp, _I,
                            _lp, _dp, _dp, _L, _dp, _dp, _dp, _dp]),
    "sep_refine": (_I, [_V, _L, _dp, _dp]),
    "sep_filter_dev": (_I, [_V, _V, _L, _dp, _dp, _L, _L, _D, _D, _I]),
    "sep_bench_hpsi": (_I, [_V, _I, _L, _L, _L, _dp]),
    "sep_bench_copy": (_I, [_V, _L, _L, _dp]),
    "sep_set_timing": (_I, [_V, _I]),
    "sep_set_filter_lanes": (_I, [_V, _I]),
    "sep_set_hpsi_family": (_I, [_V, _I, _I]),
    "sep_so_setup": (_I, [_V, _L, _dp, _dp, _lp]),
    "sep_energy_range_log": (_I, [_V, _D, _dp, _dp, _lp, _dp, _L]),
    "sep_refine_window": (_I, [_V, _L, _dp, _dp, _D, _D, _I, _lp]),
    "sep_apply_h_spinor_dev": (_I, [_V, _V, _V, _L]),
    "sep_energy_range_spinor": (_I, [_V, _dp, _dp, _lp]),
    "sep_filter_spinor_dev": (_I, [_V, _V, _L, _dp, _dp, _L, _L, _D, _D]),
    "sep_filter_spinor": (_I, [_V, _dp, _L, _dp, _dp, _L, _L, _D, _D]),
    "sep_energy_spinor": (_I, [_V, _L, _dp]),
    "sep_sigma_spinor": (_I, [_V, _L, _dp]),
    "sep_ortho_spinor": (_I, [_V, _L, _D, _dp, _lp]),
    "sep_rayleigh_ritz_spinor": (_I, [_V, _L, _dp]),
    "sep_get_hpsi_family": (_I, [_V, _I, _ip]),
    "sep_autotune_ex": (_I, [_V, _I, _L, _dp, _ip, _ip]),
    "sep_trim": (_I, [_V]),
    "sep_autotune": (_I, [_V, _I, _L, _dp, _dp]),
    "sep_filter_batch": (_I, [_I, _V, _V, _lp, _V, _V, _lp, _lp, _dp, _dp, _I, _I]),
    "sep_get_timing": (_I, [_V, _dp, ctypes.POINTER(ctypes.c_long)]),
    "sep_gram_dev": (_I, [_V, _I, _V, _L, _V, _L, _dp]),
    "sep_rotate_dev": (_I, [_V, _I, _V, _L, _dp, _L, _V]),
    "sep_deck_read": (_I, [ctypes.c_char_p, _V, ctypes.c_char_p, _I]),
    "sep_deck_free": (None, [_V]),
    "sep_type_symbol": (ctypes.c_char_p, [_I]),
}

_lib = None


def lib():
    """Load libsep.so (raises if it has not been built).

    torch, when importable, is imported first: its wheel bundles a HIP runtime
    with the same soname as ROCm's, and loading it first makes libsep and
    torch share one runtime in the process (torch is only used for device
    buffers by callers such as the autograd wrapper)."""
    global _lib
    if _lib is None:
        try:
            import torch  # noqa: F401
        except Exception:
            pass
        if not os.path.exists(LIB_PATH):
            raise SepError(f"{LIB_PATH} is not built: run __graft_entry__.build() "
                           "(make -C semi_empirical_pseudopotentials_amd)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ------------------------------------------------------------ host helpers
def ran_nrc_stream(idum, count, rng=None):
    """`count` draws of ran_nrc (filter/rand.c:28-66); returns (values, idum, rng)."""
    L = lib()
    if rng is None:
        rng = Rng()
        L.sep_rng_reset(ctypes.byref(rng))
    idp = np.array([idum], dtype=np.int64)
    out = np.empty(count)
    for i in range(count):
        out[i] = L.sep_ran_nrc(ctypes.byref(rng), idp.ctypes.data_as(_lp))
    return out, int(idp[0]), rng


def init_psi(idum, ngrid, dv, rng=None):
    """init_psi (filter/init.c:236-251); returns (psi, idum, rng)."""
    L = lib()
    if rng is None:
        rng = Rng()
        L.sep_rng_reset(ctypes.byref(rng))
    idp = np.array([idum], dtype=np.int64)
    out = np.empty(ngrid)
    L.sep_init_psi(ctypes.byref(rng), idp.ctypes.data_as(_lp), ngrid, dv, _p(out))
    return out, int(idp[0]), rng


def leja_points(nc):
    zn = np.empty(nc)
    fk = lib().sep_leja_points(nc, _p(zn))
    return zn, fk


def newton_coeffs(zn, el, dE, vmin, dt):
    zn = _f64(zn)
    el = _f64(el)
    an = np.empty((el.size, zn.size))
    rc = lib().sep_newton_coeffs(zn.size, el.size, _p(zn), _p(el), dE, vmin, dt, _p(an))
    if rc != SEP_OK:
        raise SepError(f"sep_newton_coeffs: {_ERRNAMES.get(rc, rc)}")
    return an


def read_deck(path):
    """Parse a deck directory with libsep's reader; returns a dict of numpy arrays."""
    L = lib()
    d = Deck()
    err = ctypes.create_string_buffer(512)
    rc = L.sep_deck_read(path.encode(), ctypes.byref(d), err, 512)
    if rc != SEP_OK:
        raise SepError(f"sep_deck_read: {err.value.decode()}")
    na = d.natom
    out = {k: getattr(d, k) for k in ("nx", "ny", "nz", "ms", "ns", "nc", "nthreads", "Elmin",
                                        "Elmax", "natom", "has_strain", "has_expot", "xd", "yd",
                                        "zd", "xmin", "ymin", "zmin", "dx", "dy", "dz", "dv",
                                        "seed", "has_seed")}
    out["xyz"] = np.ctypeslib.as_array(d.xyz, shape=(na, 3)).copy()
    out["type"] = np.ctypeslib.as_array(d.type, shape=(na,)).copy()
    out["scale"] = np.ctypeslib.as_array(d.scale, shape=(na,)).copy()
    out["expot"] = np.array(d.expot[:])
    out["npot"] = np.array(d.npot[:], dtype=np.int64)
    out["dr"] = np.array(d.dr[:])
    out["vr"] = np.ctypeslib.as_array(d.vr, shape=(NTYPE, NPOT)).copy()
    out["pot"] = np.ctypeslib.as_array(d.pot, shape=(NTYPE, NPOT)).copy()
    L.sep_deck_free(ctypes.byref(d))
    return out


def fft_length_supported(n):
    return bool(lib().sep_fft_length_supported(n))


SEP_MAX_LANES = 4
SEP_BATCH_INTERLEAVE = 0
SEP_BATCH_GRAPH = 1


def filter_batch(ctxs, psi0_ptrs, ns, ans, zns, vmins, dEs, precision=SEP_PREC_FP32,
                 mode=SEP_BATCH_INTERLEAVE):
    """sep_filter_batch: one filter pass on several contexts (grids) at once.

    psi0_ptrs: device pointers of each grid's ns x N FP64 random states;
    ans: (ms, nc) coefficient arrays; zns: Leja points; per-grid vmin and dE."""
    nd = len(ctxs)
    ans = [_f64(a) for a in ans]
    zns = [_f64(z) for z in zns]
    H = (_V * nd)(*[c._h for c in ctxs])
    P = (_V * nd)(*[_V(p) for p in psi0_ptrs])
    A = (_dp * nd)(*[_p(a) for a in ans])
    Z = (_dp * nd)(*[_p(z) for z in zns])
    nsv = np.ascontiguousarray(ns, dtype=np.int64)
    ncv = np.ascontiguousarray([a.shape[1] for a in ans], dtype=np.int64)
    msv = np.ascontiguousarray([a.shape[0] for a in ans], dtype=np.int64)
    vm = _f64(vmins)
    de = _f64(dEs)
    rc = lib().sep_filter_batch(nd, ctypes.cast(H, _V), ctypes.cast(P, _V),
                                nsv.ctypes.data_as(_lp), ctypes.cast(A, _V), ctypes.cast(Z, _V),
                                ncv.ctypes.data_as(_lp), msv.ctypes.data_as(_lp), _p(vm), _p(de),
                                precision, mode)
    if rc != SEP_OK:
        raise SepError(f"sep_filter_batch failed: {_ERRNAMES.get(rc, rc)}: "
                       f"{lib().sep_last_error(ctxs[0]._h).decode()}")


class Context:
    """One grid on one GPU (sep_open / sep_close)."""

    def __init__(self, nx, ny, nz, dx, dy, dz, xmin=0.0, ymin=0.0, zmin=0.0, ekinmax=10.0,
                 device=0):
        L = lib()
        self.grid = (nx, ny, nz)
        self.N = nx * ny * nz
        self.dv = dx * dy * dz
        self.device = device
        self._desc = GridDesc(nx, ny, nz, dx, dy, dz, xmin, ymin, zmin, ekinmax, device)
        h = _V()
        rc = L.sep_open(ctypes.byref(self._desc), ctypes.byref(h))
        if rc != SEP_OK:
            raise SepError(f"sep_open failed: {_ERRNAMES.get(rc, rc)}")
        self._h = h
        self.vmin = None
        self.vmax = None

    def close(self):
        if getattr(self, "_h", None):
            lib().sep_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc, what):
        if rc != SEP_OK:
            msg = lib().sep_last_error(self._h)
            raise SepError(f"{what}: {_ERRNAMES.get(rc, rc)}: {msg.decode() if msg else ''}")

    # -------------------------------------------------------- potential
    def set_potential(self, potl):
        potl = _f64(potl).reshape(-1)
        assert potl.size == self.N
        self._chk(lib().sep_set_potential(self._h, _p(potl)), "sep_set_potential")
        self.vmin, self.vmax = float(potl.min()), float(potl.max())

    def set_interp(self, interp):
        """SEP_INTERP_LINEAR (reference) or SEP_INTERP_CUBIC table interpolation for
        build_potential and hf_adjoint on this context."""
        self._chk(lib().sep_set_interp(self._h, int(interp)), "sep_set_interp")

    def build_potential(self, xyz, types, npot, vr, pot, dr, scale=None, expot=None):
        xyz = _f64(xyz).reshape(-1, 3)
        t = np.ascontiguousarray(types, dtype=np.int32)
        npot = np.ascontiguousarray(npot, dtype=np.int64)
        vr = _f64(vr)
        pot = _f64(pot)
        dr = _f64(dr)
        out = np.empty(self.N)
        vmin = np.zeros(1)
        vmax = np.zeros(1)
        sc = _p(_f64(scale)) if scale is not None else None
        ex = _p(_f64(expot)) if expot is not None else None
        self._keep = (scale if scale is None else _f64(scale))
        rc = lib().sep_build_potential(self._h, xyz.shape[0], _p(xyz), t.ctypes.data_as(_ip), sc,
                                       vr.shape[0], npot.ctypes.data_as(_lp), _p(vr), _p(pot),
                                       vr.shape[1], _p(dr), ex, _p(out), _p(vmin), _p(vmax))
        self._chk(rc, "sep_build_potential")
        self.vmin, self.vmax = float(vmin[0]), float(vmax[0])
        return out

    # ------------------------------------------------------------ H psi
    def apply_h(self, psi):
        """H on complex128 vectors shaped (batch, N) or (N,)."""
        v = np.ascontiguousarray(psi, dtype=np.complex128)
        shape = v.shape
        v = v.reshape(-1, self.N)
        out = np.empty_like(v)
        rc = lib().sep_apply_h(self._h, v.ctypes.data_as(_dp), out.ctypes.data_as(_dp), v.shape[0])
        self._chk(rc, "sep_apply_h")
        return out.reshape(shape)

    def hnorm(self, psi, zm1, vmin, dE_1):
        v = np.ascontiguousarray(psi, dtype=np.complex128)
        shape = v.shape
        v = v.reshape(-1, self.N)
        out = np.empty_like(v)
        rc = lib().sep_hnorm(self._h, v.ctypes.data_as(_dp), out.ctypes.data_as(_dp), v.shape[0],
                             zm1, vmin, dE_1)
        self._chk(rc, "sep_hnorm")
        return out.reshape(shape)

    def apply_h_dev(self, in_ptr, out_ptr, batch, precision=SEP_PREC_FP64):
        rc = lib().sep_apply_h_dev(self._h, _V(in_ptr), _V(out_ptr), batch, precision)
        self._chk(rc, "sep_apply_h_dev")

    def energy_range(self, vmin=None):
        vmin = self.vmin if vmin is None else vmin
        dE = np.zeros(1)
        emax = np.zeros(1)
        it = np.zeros(1, dtype=np.int64)
        rc = lib().sep_energy_range(self._h, vmin, _p(dE), _p(emax), it.ctypes.data_as(_lp))
        self._chk(rc, "sep_energy_range")
        return float(dE[0]), float(emax[0]), int(it[0])

    # ------------------------------------------------------------ filter
    def filter(self, psi0, an, zn, vmin, dE, precision=SEP_PREC_FP64, want_states=True,
               want_energies=True):
        psi0 = _f64(psi0).reshape(-1, self.N)
        an = _f64(an)
        zn = _f64(zn)
        ns = psi0.shape[0]
        ms, nc = an.shape
        out = np.empty((ns * ms, self.N)) if want_states else None
        ene = np.empty(ns * ms) if want_energies else None
        rc = lib().sep_filter(self._h, _p(psi0), ns, _p(an), _p(zn), nc, ms, vmin, dE, precision,
                              _p(out) if out is not None else None,
                              _p(ene) if ene is not None else None)
        self._chk(rc, "sep_filter")
        return out, ene

    def filter_dev(self, psi0_ptr, ns, an, zn, vmin, dE, precision=SEP_PREC_FP64):
        """Filter with the random states already on the device (ns x N FP64)."""
        an = _f64(an)
        zn = _f64(zn)
        ms, nc = an.shape
        rc = lib().sep_filter_dev(self._h, _V(psi0_ptr), ns, _p(an), _p(zn), nc, ms, vmin, dE,
                                  precision)
        self._chk(rc, "sep_filter_dev")

    def bench_hpsi(self, precision, nvec, ms, iters):
        out = np.zeros(6)
        rc = lib().sep_bench_hpsi(self._h, precision, nvec, ms, iters, _p(out))
        self._chk(rc, "sep_bench_hpsi")
        return out

    def autotune(self, precision, nvec=8):
        """Pick the faster H.psi kernel family for this grid (sep_autotune)."""
        tp, tl = np.zeros(1), np.zeros(1)
        self._chk(lib().sep_autotune(self._h, precision, nvec, _p(tp), _p(tl)), "sep_autotune")
        return float(tp[0]), float(tl[0])

    def set_hpsi_family(self, precision, family):
        """H.psi kernel family (SEP_HPSI_PLANE / _PLANE_CHUNKED / _LINES / _PLANE_SWIZZLED)
        for this grid."""
        self._chk(lib().sep_set_hpsi_family(self._h, precision, family), "sep_set_hpsi_family")

    def hpsi_family(self, precision):
        f = ctypes.c_int(0)
        self._chk(lib().sep_get_hpsi_family(self._h, precision, ctypes.byref(f)),
                  "sep_get_hpsi_family")
        return f.value

    def autotune_ex(self, precision, nvec=8):
        """Fastest (family, lanes) by short timed filters (sep_autotune_ex);
        returns (ms per Newton step per family, -1 where absent), family, lanes."""
        ms = np.zeros(SEP_HPSI_NFAMILY)
        fam, lanes = ctypes.c_int(0), ctypes.c_int(0)
        self._chk(lib().sep_autotune_ex(self._h, precision, nvec, _p(ms), ctypes.byref(fam),
                                        ctypes.byref(lanes)), "sep_autotune_ex")
        return ms, fam.value, lanes.value

    # ---- spin-orbit spinors (FP64; a spinor is 4 rows of N: up.re, up.im, dn.re, dn.im)
    def so_setup(self, xyz, vso):
        """Spin-orbit neighbourhood lists (sep_so_setup); returns the entry count."""
        x = np.ascontiguousarray(xyz, dtype=np.float64)
        v = np.ascontiguousarray(vso, dtype=np.float64)
        n = ctypes.c_long(0)
        self._chk(lib().sep_so_setup(self._h, len(v), _p(x), _p(v), ctypes.byref(n)),
                  "sep_so_setup")
        return n.value

    def apply_h_spinor_dev(self, in_ptr, out_ptr, nspinor):
        self._chk(lib().sep_apply_h_spinor_dev(self._h, ctypes.c_void_p(in_ptr),
                                               ctypes.c_void_p(out_ptr), nspinor),
                  "sep_apply_h_spinor_dev")

    def energy_range_spinor(self):
        """(Vmin, dE, (Emin iterations, Emax iterations)) (sep_energy_range_spinor)."""
        vmin, dE = ctypes.c_double(0.0), ctypes.c_double(0.0)
        its = np.zeros(2, dtype=np.int64)
        self._chk(lib().sep_energy_range_spinor(self._h, ctypes.byref(vmin), ctypes.byref(dE),
                                                its.ctypes.data_as(_lp)), "sep_energy_range_spinor")
        return vmin.value, dE.value, (int(its[0]), int(its[1]))

    def filter_spinor_dev(self, psi0_ptr, nspinor, an, zn, vmin, dE):
        an = np.ascontiguousarray(an, dtype=np.float64)
        zn = np.ascontiguousarray(zn, dtype=np.float64)
        self._chk(lib().sep_filter_spinor_dev(self._h, ctypes.c_void_p(psi0_ptr), nspinor, _p(an),
                                              _p(zn), zn.size, an.shape[0], vmin, dE),
                  "sep_filter_spinor_dev")

    def energy_spinor(self, m):
        e = np.zeros(m)
        self._chk(lib().sep_energy_spinor(self._h, m, _p(e)), "sep_energy_spinor")
        return e

    def sigma_spinor(self, m):
        s = np.zeros(m)
        self._chk(lib().sep_sigma_spinor(self._h, m, _p(s)), "sep_sigma_spinor")
        return s

    def ortho_spinor(self, m, cutoff=1.0e-10):
        S = np.zeros(m)
        k = ctypes.c_long(0)
        self._chk(lib().sep_ortho_spinor(self._h, m, cutoff, _p(S), ctypes.byref(k)),
                  "sep_ortho_spinor")
        return k.value, S

    def rayleigh_ritz_spinor(self, m):
        e = np.zeros(m)
        self._chk(lib().sep_rayleigh_ritz_spinor(self._h, m, _p(e)), "sep_rayleigh_ritz_spinor")
        return e

    def spinors_download(self, m):
        """The first m resident spinors as complex (m, 2, N) (up, down)."""
        x = self.states_download(4 * m).reshape(m, 2, 2, -1)
        return x[:, :, 0, :] + 1j * x[:, :, 1, :]

    def set_timing(self, on):
        """Sample kernel durations inside the filter loop (see sep_set_timing)."""
        self._chk(lib().sep_set_timing(self._h, 1 if on else 0), "sep_set_timing")

    def trim(self):
        """Free the filter's ring of Newton vectors (sep_trim)."""
        self._chk(lib().sep_trim(self._h), "sep_trim")

    def set_filter_lanes(self, lanes):
        """Concurrent groups of random states in the filter (sep_set_filter_lanes):
        0 = automatic, 1 = off, up to SEP_MAX_LANES."""
        self._chk(lib().sep_set_filter_lanes(self._h, int(lanes)), "sep_set_filter_lanes")

    def get_timing(self):
        """(avg ms of passes A, B, C, avg ms of one full-block accumulation), samples."""
        out = np.zeros(4)
        n = (ctypes.c_long * 2)()
        self._chk(lib().sep_get_timing(self._h, _p(out), n), "sep_get_timing")
        return out, (n[0], n[1])

    def bench_copy(self, nbytes=4 << 30, iters=5):
        """Achievable HBM bandwidth, GB/s read + written (sep_bench_copy)."""
        out = np.zeros(1)
        self._chk(lib().sep_bench_copy(self._h, int(nbytes), int(iters), _p(out)), "sep_bench_copy")
        return float(out[0])

    def gram_dev(self, A_ptr, m, B_ptr, mb, precision=SEP_PREC_FP64):
        C = np.empty((m, mb))
        rc = lib().sep_gram_dev(self._h, precision, _V(A_ptr), m, _V(B_ptr), mb, _p(C))
        self._chk(rc, "sep_gram_dev")
        return C

    def rotate_dev(self, A_ptr, m, M, out_ptr, precision=SEP_PREC_FP64):
        M = _f64(M)
        rc = lib().sep_rotate_dev(self._h, precision, _V(A_ptr), m, _p(M), M.shape[1],
                                  _V(out_ptr))
        self._chk(rc, "sep_rotate_dev")

    # ------------------------------------------------------------ states
    def states_upload(self, psi):
        psi = _f64(psi).reshape(-1, self.N)
        self._chk(lib().sep_states_upload(self._h, _p(psi), psi.shape[0]), "sep_states_upload")
        return psi.shape[0]

    def states_download(self, m):
        out = np.empty((m, self.N))
        self._chk(lib().sep_states_download(self._h, _p(out), m), "sep_states_download")
        return out

    def normalize_all(self, m):
        self._chk(lib().sep_normalize_all(self._h, m), "sep_normalize_all")

    def energy_all(self, m):
        e = np.empty(m)
        self._chk(lib().sep_energy_all(self._h, m, _p(e)), "sep_energy_all")
        return e

    def ortho(self, m, cutoff=1.0e-10):
        S = np.empty(m)
        kept = np.zeros(1, dtype=np.int64)
        rc = lib().sep_ortho(self._h, m, cutoff, _p(S), kept.ctypes.data_as(_lp))
        self._chk(rc, "sep_ortho")
        return int(kept[0]), S

    def rayleigh_ritz(self, m):
        ev = np.empty(m)
        self._chk(lib().sep_rayleigh_ritz(self._h, m, _p(ev)), "sep_rayleigh_ritz")
        return ev

    def sigma(self, m):
        s = np.empty(m)
        self._chk(lib().sep_sigma(self._h, m, _p(s)), "sep_sigma")
        return s

    def hf_adjoint(self, pairs, xyz, types, npot, vr, pot, dr, scale=None, weights=None):
        """Hellmann-Feynman contractions of resident-state densities (sep_hf_adjoint).

        pairs: (P, 2) resident-state indices; weights (P) combine them into one
        density.  Returns (dR [P', natom, 3], dscale [P', natom],
        dtab [P', ntype, stride]) with P' = 1 if weights is given, else P."""
        pairs = np.ascontiguousarray(np.asarray(pairs, dtype=np.int64).reshape(-1, 2))
        npairs = pairs.shape[0]
        xyz = _f64(xyz)
        natom = xyz.shape[0]
        t = np.ascontiguousarray(np.asarray(types, dtype=np.int32))
        vr = _f64(vr)
        pot = _f64(pot)
        ntype, stride = vr.shape
        npot = np.ascontiguousarray(np.asarray(npot, dtype=np.int64))
        nout = 1 if weights is not None else npairs
        dR = np.zeros((nout, natom, 3))
        ds = np.zeros((nout, natom))
        dtab = np.zeros((nout, ntype, stride))
        w = _f64(weights) if weights is not None else None
        sc = _f64(scale) if scale is not None else None
        rc = lib().sep_hf_adjoint(self._h, npairs, pairs.ctypes.data_as(_lp),
                                  _p(w) if w is not None else None, natom, _p(xyz),
                                  t.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                  _p(sc) if sc is not None else None, ntype,
                                  npot.ctypes.data_as(_lp), _p(vr), _p(pot), stride, _p(_f64(dr)),
                                  _p(dR), _p(ds), _p(dtab))
        self._chk(rc, "sep_hf_adjoint")
        return dR, ds, dtab

    def refine_window(self, ev, sg, lo, hi, passes=2, full=False):
        """sep_refine_window: more correction passes on the resident Ritz pairs
        with energies in [lo, hi], refined in place.  Returns the (ev, sigma)
        of the pairs in [lo, hi] after the passes, or with full=True those of
        all len(ev) resident pairs (ascending)."""
        ev = np.array(ev, dtype=np.float64)
        sg = np.array(sg, dtype=np.float64)
        kw = ctypes.c_long(0)
        self._chk(lib().sep_refine_window(self._h, ev.size, _p(ev), _p(sg), lo, hi, passes,
                                          ctypes.byref(kw)), "sep_refine_window")
        if full:
            return ev, sg
        inw = (ev >= lo) & (ev <= hi)
        return ev[inw].copy(), sg[inw].copy()

    def ortho_fp32(self, m, cutoff=1.0e-10):
        """ortho of the FP32 pipeline (sep_ortho_fp32): FP64 Gram eigenbasis with the
        SEP_FP32_SVD_FLOOR floor, polished by Cholesky-QR."""
        S = np.empty(m)
        kept = np.zeros(1, dtype=np.int64)
        rc = lib().sep_ortho_fp32(self._h, m, cutoff, _p(S), kept.ctypes.data_as(_lp))
        self._chk(rc, "sep_ortho_fp32")
        return int(kept[0]), S

    def refine(self, ev):
        """One FP64 kinetically preconditioned correction pass on the resident Ritz pairs."""
        ev = np.array(ev, dtype=np.float64)
        sg = np.empty_like(ev)
        self._chk(lib().sep_refine(self._h, ev.size, _p(ev), _p(sg)), "sep_refine")
        return ev, sg
