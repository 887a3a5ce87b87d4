"""The `filter` driver (drop-in for the reference's filter.x) end to end:
same eval.dat/psi.dat formats; levels match the CPU oracle run on the same
seed, and the converged levels of the reference's own In13_P16 run
(filterStrainCube/test/In13_P16/eval.dat) to 1e-8 Ha."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from tests.conftest import GOLDEN
from tests.deckutil import match_levels, read_eval

pytestmark = pytest.mark.gpu


def _run_filter(deck, out, *args):
    from semi_empirical_pseudopotentials_amd._lib import FILTER_BIN
    shutil.copytree(deck, out, dirs_exist_ok=True)
    r = subprocess.run([FILTER_BIN, "--dir", out, *args], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout


def test_driver_small_vs_oracle(sep, oracle, small_deck, tmp_path):
    print("case driver small", flush=True)
    out = str(tmp_path / "gpu")
    log = _run_filter(small_deck, out, "--seed", "-1001")
    s = oracle.run(small_deck, str(tmp_path / "cpu"), -1001, 0)
    ge, gs = read_eval(os.path.join(out, "eval.dat"))
    oe, os_ = read_eval(str(tmp_path / "cpu" / "eval.dat"))
    assert ge.size == oe.size == int(s["cutoff"])
    good = os_ <= np.median(os_)
    assert np.abs(ge[good] - oe[good]).max() < 1e-8
    N = 24 ** 3
    assert os.path.getsize(os.path.join(out, "psi.dat")) == 8 * N * ge.size
    assert "cutoff is" in log and "dE =" in log
    for j in range(3):
        assert os.path.exists(os.path.join(out, f"eval-filt-0-{j}.dat"))
    _check_side_files(sep, out, ge.size)


def _check_side_files(sep, out, kept):
    """conf.dat, pot<Sym>.dat, pS.dat, localPot.cube, ene-init.dat,
    prop0.dat and hmat.dat in the reference's formats (filter/read.c:32,114,
    init.c:185-192, write.c:31-119, energy.c:65, filter.c:69, Hmat.c:28)."""
    from semi_empirical_pseudopotentials_amd._lib import read_deck
    d = read_deck(out)
    conf = [l.split() for l in open(os.path.join(out, "conf.dat")).read().splitlines()]
    assert len(conf) == d["natom"] and all(len(c) == 5 for c in conf)
    np.testing.assert_allclose(np.array([[float(v) for v in c[1:4]] for c in conf]), d["xyz"],
                               rtol=1e-5, atol=1e-5)
    cd = np.loadtxt(os.path.join(out, "potCd.dat"))
    assert cd.shape[0] == d["npot"][0]
    np.testing.assert_allclose(cd[:, 1], d["pot"][0][:cd.shape[0]], rtol=1e-5, atol=1e-9)
    ps = np.loadtxt(os.path.join(out, "pS.dat"))
    assert ps.shape == (d["nz"], 2)
    lines = open(os.path.join(out, "localPot.cube")).read().splitlines()
    assert lines[0] == "CUBE FILE" and int(lines[2].split()[0]) == d["natom"]
    vals = np.array(" ".join(lines[6 + d["natom"]:]).split(), dtype=float)
    N = d["nx"] * d["ny"] * d["nz"]
    assert vals.size == N
    with sep.Context(d["nx"], d["ny"], d["nz"], d["dx"], d["dy"], d["dz"], d["xmin"], d["ymin"],
                     d["zmin"]) as ctx:
        V = ctx.build_potential(d["xyz"], d["type"], d["npot"], d["vr"], d["pot"], d["dr"],
                                expot=d["expot"])
    np.testing.assert_allclose(vals, V, rtol=1e-5, atol=1e-9)  # "%g": 6 digits
    ene = np.loadtxt(os.path.join(out, "ene-init.dat"), ndmin=2)
    assert ene.shape[1] == 4 and np.array_equal(ene[:, 0], np.arange(ene.shape[0]))
    assert np.array_equal(ene[1:, 1], ene[:-1, 2])  # each line's old energy is the last one's
    hm = open(os.path.join(out, "hmat.dat")).read().splitlines()
    assert hm[0] == "calculate the H matrix" and hm[-1] == "free memory"
    assert sum(l.startswith("finshed row") for l in hm) == kept
    assert os.path.exists(os.path.join(out, "prop0.dat"))


def test_driver_in13_reference(sep, tmp_path):
    print("case driver In13_P16", flush=True)
    out = str(tmp_path / "in13")
    log = _run_filter(os.path.join(GOLDEN, "In13_P16"), out, "--seed", "-1131359211")
    assert "dE = 13.3627" in log
    ref_e, ref_s = read_eval(os.path.join(GOLDEN, "In13_P16", "eval.dat"))
    got_e, got_s = read_eval(os.path.join(out, "eval.dat"))
    assert got_e.size == 158
    pairs = match_levels(ref_e, ref_s, got_e, 1e-6)
    assert len(pairs) >= 40
    assert np.abs(pairs[:, 0] - pairs[:, 1]).max() < 1e-8


def test_driver_in13_fp32_window_vs_reference(sep, tmp_path):
    """--precision fp32 on In13_P16: complex64 filter and ortho, FP64
    Rayleigh-Ritz and sigma, then 10 FP64 correction passes on the pairs in
    [Elmin, Elmax] of input.par (-0.23, -0.05 Ha).  Every level of the
    reference's own eval.dat (FP64) converged there (sigma < 1e-4) is
    matched one to one, multiplet by multiplet, within 1 meV."""
    from tests.deckutil import compare_levels
    print("case driver In13_P16 fp32", flush=True)
    out = str(tmp_path / "in13_32")
    log = _run_filter(os.path.join(GOLDEN, "In13_P16"), out, "--seed", "-1131359211",
                      "--precision", "fp32")
    assert "FP64 correction passes" in log
    ref_e, ref_s = read_eval(os.path.join(GOLDEN, "In13_P16", "eval.dat"))
    got_e, got_s = read_eval(os.path.join(out, "eval.dat"))
    assert np.all(np.diff(got_e) >= 0)
    N = 36 ** 3
    assert os.path.getsize(os.path.join(out, "psi.dat")) == 8 * N * got_e.size
    mev = 1e-3 / 27.211386245988
    nc, nl, d = compare_levels(ref_e, ref_s, got_e, got_s, -0.23, -0.05, mev)
    print(f"fp32 driver vs reference eval.dat in [-0.23, -0.05]: {nl} levels in {nc} "
          f"multiplets, max |dE| {d / mev:.4f} meV", flush=True)
    # every converged reference level of the window: 7 (3 multiplets)
    inwin = (ref_s < 1e-4) & (ref_e >= -0.23) & (ref_e <= -0.05)
    assert nl == inwin.sum() == 7


def test_driver_resume_from_filtered_states(sep, small_deck, tmp_path):
    """--psi-filt writes the filtered states; --resume-filt in a copy of that
    directory skips the filter and reproduces eval.dat and psi.dat exactly."""
    print("case driver resume", flush=True)
    first = str(tmp_path / "first")
    _run_filter(small_deck, first, "--seed", "-1001", "--psi-filt")
    again = str(tmp_path / "again")
    shutil.copytree(first, again)
    for f in ("eval.dat", "psi.dat"):
        os.remove(os.path.join(again, f))
    from semi_empirical_pseudopotentials_amd._lib import FILTER_BIN
    r = subprocess.run([FILTER_BIN, "--dir", again, "--resume-filt"], capture_output=True,
                       text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "filtered states from psi-filt" in r.stdout
    assert "done calculating filter" not in r.stdout
    for f in ("eval.dat", "psi.dat"):
        assert open(os.path.join(first, f), "rb").read() == open(os.path.join(again, f), "rb").read()
    os.remove(os.path.join(again, "psi-filt-0-1.dat"))
    r = subprocess.run([FILTER_BIN, "--dir", again, "--resume-filt"], capture_output=True,
                       text=True, timeout=600)
    assert r.returncode == 1
    assert "psi-filt" in open(os.path.join(again, "error")).read()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_driver_lanes_same_bytes(sep, small_deck, tmp_path, prec):
    """--lanes 2 (two groups of random states on concurrent streams) writes
    eval.dat and psi.dat byte for byte as --lanes 1."""
    print("case driver lanes", prec, flush=True)
    outs = []
    for lanes in ("1", "2"):
        out = str(tmp_path / f"l{lanes}")
        _run_filter(small_deck, out, "--seed", "-1001", "--precision", prec, "--lanes", lanes)
        outs.append(out)
    for f in ("eval.dat", "psi.dat"):
        a = open(os.path.join(outs[0], f), "rb").read()
        b = open(os.path.join(outs[1], f), "rb").read()
        assert len(a) > 0 and a == b, f
