"""The GPU filter-diagonalization pipeline against the CPU oracle on the same
seeded inputs: V(r) (filter/init.c:157-199), the spectral bound
(filter/energy.c:53-83), the Newton filter (filter/filter.c:59-83), ortho
(filter/ortho.c:11-52), Rayleigh-Ritz (filter/Hmat.c:11-86) and sigma
(filter/energy.c:88-119).  FP64 tolerances are written in each test."""
import numpy as np
import pytest

from tests.deckutil import context_for, device_potential, oracle_potential

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(sep, oracle, small_deck):
    from semi_empirical_pseudopotentials_amd._lib import read_deck
    d = read_deck(small_deck)
    grid = (d["nx"], d["ny"], d["nz"])
    V = oracle_potential(oracle, d)
    ksq = oracle.ksqr(*grid, d["dx"], d["dy"], d["dz"])
    return d, grid, V, ksq


def test_potential_device(sep, oracle, setup):
    d, grid, V, ksq = setup
    print("case potential", grid, flush=True)
    with context_for(sep, d) as ctx:
        got = device_potential(ctx, d)
        assert np.abs(got - V).max() <= 1e-14 * np.abs(V).max()
        assert ctx.vmin == pytest.approx(V.min(), abs=1e-14)


def test_energy_range(sep, oracle, setup):
    d, grid, V, ksq = setup
    print("case energy_range", grid, flush=True)
    dE_o, emax_o, it_o = oracle.energy_range(grid, ksq, V, d["dv"], V.min())
    with context_for(sep, d) as ctx:
        ctx.set_potential(V)
        dE, emax, it = ctx.energy_range()
    assert it == it_o
    assert abs(dE - dE_o) < 1e-10 * dE_o


def _coeffs(oracle, d, V, ksq, grid, nc, ms):
    dE, _, _ = oracle.energy_range(grid, ksq, V, d["dv"], V.min())
    zn, _ = oracle.leja_points(nc)
    el = np.linspace(d["Elmin"], d["Elmax"], ms)
    dt = (nc / (2.5 * dE)) ** 2
    an = oracle.newton_coeffs(nc, el, zn, dE, V.min(), dt)
    return dE, zn, el, an


def _random_states(oracle, ns, N, dv, seed=-1001):
    r = oracle.Rng()
    idc = np.array([seed], dtype=np.int64)
    return np.array([r.init_psi(idc, N, dv) for _ in range(ns)])


@pytest.mark.parametrize("ns,depth", [(3, None), (4, None), (3, "32"), (4, "40")])
def test_filter_fp64(sep, oracle, setup, monkeypatch, ns, depth):
    """Filtered states (normalised) match the oracle's to 1e-9 relative; with
    the ring depth bounded (test hook SEP_RING_DEPTH) the 95 Newton vectors
    are accumulated over 3 ring blocks, the last one partial."""
    if depth:
        monkeypatch.setenv("SEP_RING_DEPTH", depth)
    d, grid, V, ksq = setup
    print("case filter_fp64 ns", ns, "ring depth", depth, flush=True)
    N, nc, ms = V.size, 96, 4
    dE, zn, el, an = _coeffs(oracle, d, V, ksq, grid, nc, ms)
    psi0 = _random_states(oracle, ns, N, d["dv"])
    ref = []
    for j in range(ns):
        f = oracle.filter_state(grid, ksq, V, an, zn, 4.0 / dE, V.min(), psi0[j])
        ref.append(oracle.normalize_all(f, d["dv"]))
    ref = np.concatenate(ref)
    ene_ref = oracle.energy_all(grid, ksq, V, d["dv"], ref)
    with context_for(sep, d) as ctx:
        ctx.set_potential(V)
        got, ene = ctx.filter(psi0, an, zn, V.min(), dE, sep.SEP_PREC_FP64)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 1e-9, err
    assert np.abs(ene - ene_ref).max() < 1e-10


@pytest.mark.parametrize("nc,ms,depth", [(96, 4, None), (37, 5, None), (13, 33, None),
                                         (96, 4, "32"), (96, 5, "24")])
def test_filter_fp32(sep, oracle, setup, monkeypatch, nc, ms, depth):
    """complex64 recurrence with FP64 accumulators: close to the FP64 oracle.
    Newton lengths that are not a multiple of 16 (the packed accumulation's
    partial last row group), more than 32 targets (a second target group
    of one) and bounded ring depths (several ring blocks; 24 is not a
    multiple of 16, so the general accumulation kernel runs) included."""
    if depth:
        monkeypatch.setenv("SEP_RING_DEPTH", depth)
    d, grid, V, ksq = setup
    print("case filter_fp32", nc, ms, "ring depth", depth, flush=True)
    N, ns = V.size, 3
    dE, zn, el, an = _coeffs(oracle, d, V, ksq, grid, nc, ms)
    psi0 = _random_states(oracle, ns, N, d["dv"])
    ref = np.concatenate([oracle.normalize_all(
        oracle.filter_state(grid, ksq, V, an, zn, 4.0 / dE, V.min(), psi0[j]), d["dv"])
        for j in range(ns)])
    with context_for(sep, d) as ctx:
        ctx.set_potential(V)
        got, ene = ctx.filter(psi0, an, zn, V.min(), dE, sep.SEP_PREC_FP32)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 2e-4, err


def test_subspace_fp64(sep, oracle, setup):
    """ortho -> Rayleigh-Ritz -> sigma on the same filtered block: kept count
    and singular values as the oracle's SVD, Ritz values to 1e-10 Ha."""
    d, grid, V, ksq = setup
    print("case subspace", grid, flush=True)
    N, nc, ms, ns = V.size, 96, 4, 3
    dE, zn, el, an = _coeffs(oracle, d, V, ksq, grid, nc, ms)
    psi0 = _random_states(oracle, ns, N, d["dv"])
    blk = np.concatenate([oracle.normalize_all(
        oracle.filter_state(grid, ksq, V, an, zn, 4.0 / dE, V.min(), psi0[j]), d["dv"])
        for j in range(ns)])
    U, S_o, cut_o = oracle.ortho(blk)
    U = oracle.normalize_all(U[:cut_o], d["dv"])
    ev_o, rot_o = oracle.hmat(grid, ksq, V, d["dv"], U)
    rot_o = oracle.normalize_all(rot_o, d["dv"])
    sig_o = oracle.sigma(grid, ksq, V, d["dv"], rot_o)
    with context_for(sep, d) as ctx:
        ctx.set_potential(V)
        ctx.states_upload(blk)
        kept, S = ctx.ortho(blk.shape[0])
        assert kept == cut_o
        assert np.abs(S - S_o).max() < 1e-13 * S_o[0]
        Q = ctx.states_download(kept)
        assert np.abs(Q @ Q.T * d["dv"] - np.eye(kept)).max() < 1e-12
        ev = ctx.rayleigh_ritz(kept)
        sig = ctx.sigma(kept)
        # Ritz pairs on the oracle's own basis (isolates the RR stage)
        ctx.states_upload(U)
        ev2 = ctx.rayleigh_ritz(cut_o)
    assert np.abs(ev2 - ev_o).max() < 1e-10
    # same span: Ritz values of the better-converged half agree
    good = sig_o <= np.median(sig_o)
    assert np.abs(ev[good] - ev_o[good]).max() < 1e-8
    assert np.abs(sig[good] - sig_o[good]).max() < 1e-6
