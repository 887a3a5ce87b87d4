"""Random states (filter/rand.c:28-66, filter/init.c:236-251) and Newton
coefficients (filter/coeff.c:16-134): oracle pinned by an independent numpy
restatement and mpmath, product host code checked bitwise against the oracle."""
import numpy as np
import pytest


def _ran_py(state, idum):
    """Integer restatement of ran_nrc with C truncating division."""
    IM1, IM2, IA1, IA2 = 2147483563, 2147483399, 40014, 40692
    IQ1, IQ2, IR1, IR2, NTAB = 53668, 52774, 12211, 3791, 32
    IMM1 = IM1 - 1
    NDIV = 1 + IMM1 // NTAB

    def cdiv(a, b):
        q = abs(a) // abs(b)
        return q if (a >= 0) == (b > 0) else -q

    if idum[0] <= 0:
        idum[0] = 1 if -idum[0] < 1 else -idum[0]
        state["idum2"] = idum[0]
        for j in range(NTAB + 7, -1, -1):
            k = cdiv(idum[0], IQ1)
            idum[0] = IA1 * (idum[0] - k * IQ1) - k * IR1
            if idum[0] < 0:
                idum[0] += IM1
            if j < NTAB:
                state["iv"][j] = idum[0]
        state["iy"] = state["iv"][0]
    k = cdiv(idum[0], IQ1)
    idum[0] = IA1 * (idum[0] - k * IQ1) - k * IR1
    if idum[0] < 0:
        idum[0] += IM1
    k = cdiv(state["idum2"], IQ2)
    state["idum2"] = IA2 * (state["idum2"] - k * IQ2) - k * IR2
    if state["idum2"] < 0:
        state["idum2"] += IM2
    j = cdiv(state["iy"], NDIV)
    state["iy"] = state["iv"][j] - state["idum2"]
    state["iv"][j] = idum[0]
    if state["iy"] < 1:
        state["iy"] += IMM1
    t = (1.0 / IM1) * state["iy"]
    return 1.0 - 1.2e-7 if t > 1.0 - 1.2e-7 else t


@pytest.mark.parametrize("seed", [-874917403, -1, 0, -1131359211, -1001])
def test_ran_nrc_oracle_vs_restatement(oracle, seed):
    st = {"idum2": 123456789, "iy": 0, "iv": [0] * 32}
    idp = [seed]
    r = oracle.Rng()
    idc = np.array([seed], dtype=np.int64)
    for _ in range(3000):
        assert _ran_py(st, idp) == r.ran(idc)
        assert idp[0] == idc[0]


def test_ran_nrc_product_bitwise(oracle, sep):
    vals, idum, _ = sep.ran_nrc_stream(-1131359211, 5000)
    r = oracle.Rng()
    idc = np.array([-1131359211], dtype=np.int64)
    ref = np.array([r.ran(idc) for _ in range(5000)])
    assert np.array_equal(vals, ref) and idum == idc[0]


def test_init_psi_product_bitwise(oracle, sep):
    """Sequential states from one stream, as filter/main.c:82-89 draws them."""
    N, dv = 3000, 0.47
    r = oracle.Rng()
    idc = np.array([-1001], dtype=np.int64)
    ref = [r.init_psi(idc, N, dv) for _ in range(3)]
    got = []
    idum, rng = -1001, None
    for _ in range(3):
        p, idum, rng = sep.init_psi(idum, N, dv, rng)
        got.append(p)
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)
        assert abs(np.sum(a * a) * dv - 1.0) < 1e-13


def _leja_brute(nc):
    nc3 = 32 * nc
    imfrac = nc3 // 8
    nreal = nc3 - imfrac
    c = np.zeros(nc3)
    c[:nreal] = -2.0 + np.arange(nreal) * (4.0 / (nreal - 1))
    pts = [c[0]]
    for _ in range(1, nc):
        best, bv = None, -np.inf
        for k in range(nc3):
            d2 = [(c[k] - p) ** 2 for p in pts]
            if min(d2) < 1e-10:
                continue
            s = sum(np.log(d2))
            if s > bv:
                bv, best = s, k
        pts.append(c[best])
    return np.array(pts)


@pytest.mark.parametrize("nc", [4, 8, 16])
def test_leja_brute_force(oracle, nc):
    zn, _ = oracle.leja_points(nc)
    assert np.array_equal(zn, _leja_brute(nc))


def test_leja_properties(oracle):
    zn, fk = oracle.leja_points(256)
    assert zn[0] == -2.0 and zn[1] == 2.0 and zn[2] == 0.0
    assert np.all(np.abs(zn) <= 2.0) and np.unique(zn).size == zn.size
    # the origin is a node, so the capacity estimate prod|z|^(1/nc) of
    # coeff.c:295-297 is 0 (the reference never uses it)
    assert fk == 0.0
    # Leja nodes fill [-2, 2] like the equilibrium (arcsine) measure
    hist, _ = np.histogram(zn, bins=8, range=(-2, 2))
    assert hist[0] > hist[3] and hist[-1] > hist[4]


@pytest.mark.parametrize("nc", [16, 128, 1024])
def test_leja_product_bitwise(oracle, sep, nc):
    a, fa = oracle.leja_points(nc)
    b, fb = sep.leja_points(nc)
    assert np.array_equal(a, b) and fa == fb


def _coef_setup(nc):
    from oracle import oracle as O
    zn, _ = O.leja_points(nc)
    el = np.array([-0.3, -0.2, -0.1])
    dE, vmin = 13.0, -1.4
    dt = (nc / (2.5 * dE)) ** 2
    return zn, el, dE, vmin, dt


def test_newton_interpolation_identity(oracle):
    """p(z_k) = f(z_k) at every node (the defining property of coeff.c:29-53)."""
    nc = 48
    zn, el, dE, vmin, dt = _coef_setup(nc)
    an = oracle.newton_coeffs(nc, el, zn, dE, vmin, dt)
    for ie, E in enumerate(el):
        f = np.sqrt(dt / np.pi) * np.exp(-dt * ((zn + 2) * dE / 4 + vmin - E) ** 2)
        p = np.zeros(nc)
        for k in range(nc):
            w, s = 1.0, an[ie, 0]
            for j in range(1, k + 1):
                w *= zn[k] - zn[j - 1]
                s += an[ie, j] * w
            p[k] = s
        assert np.abs(p - f).max() <= 1e-9 * np.abs(f).max()


def test_newton_vs_mpmath(oracle):
    import mpmath as mp
    mp.mp.dps = 50
    nc = 24
    zn, el, dE, vmin, dt = _coef_setup(nc)
    an = oracle.newton_coeffs(nc, el, zn, dE, vmin, dt)
    E = el[1]
    z = [mp.mpf(v) for v in zn]
    f = [mp.sqrt(mp.mpf(dt) / mp.pi) * mp.exp(-mp.mpf(dt) * ((zi + 2) * mp.mpf(dE) / 4 +
                                                              mp.mpf(vmin) - mp.mpf(E)) ** 2)
         for zi in z]
    dd = list(f)
    coef = [dd[0]]
    for lvl in range(1, nc):
        dd = [(dd[i + 1] - dd[i]) / (z[i + lvl] - z[i]) for i in range(len(dd) - 1)]
        coef.append(dd[0])
    ref = np.array([float(c) for c in coef])
    scale = np.abs(ref).max()
    assert np.abs(an[1] - ref).max() <= 1e-8 * scale


@pytest.mark.parametrize("nc", [32, 512])
def test_newton_product_bitwise(oracle, sep, nc):
    zn, el, dE, vmin, dt = _coef_setup(nc)
    a = oracle.newton_coeffs(nc, el, zn, dE, vmin, dt)
    b = sep.newton_coeffs(zn, el, dE, vmin, dt)
    assert np.array_equal(a, b)
