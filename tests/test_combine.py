"""Tooling around the filter's outputs, on the CPU: merging filter runs into
one eval.par / psi.par set (combine.py, after the reference's
combine_filter_outputs.py) -- sort order, renumbering, block mapping, text
format and every refusal, on small synthetic runs -- and the couplings
driver's refusal of malformed options."""
import os

import numpy as np
import pytest

from semi_empirical_pseudopotentials_amd import combine as C


def make_run(path, energies, sigmas, ngrid, seed, conf="2\nCd 0 0 0\nSe 1 1 1\n", nxyz=None):
    os.makedirs(path, exist_ok=True)
    with open(os.path.join(path, "eval.dat"), "w") as f:
        for i, (e, s) in enumerate(zip(energies, sigmas)):
            f.write("%ld %.16g %g\n" % (i, e, s))  # the filter driver's format
    psi = np.random.default_rng(seed).standard_normal((len(energies), ngrid))
    psi.tofile(os.path.join(path, "psi.dat"))
    if conf is not None:
        with open(os.path.join(path, "conf.par"), "w") as f:
            f.write(conf)
    if nxyz is not None:
        with open(os.path.join(path, "input.par"), "w") as f:
            f.write("%d %d %d 4 2 128 -0.1 0.1 8\n" % nxyz)
    return psi


def test_merge_sorts_and_maps_blocks(tmp_path):
    ng = 24
    a = make_run(tmp_path / "a", [0.30, -0.10, 0.05], [1e-3, 2e-4, 5e-5], ng, 1, nxyz=(2, 3, 4))
    b = make_run(tmp_path / "b", [0.05, -0.20], [1e-5, 3e-4], ng, 2, nxyz=(2, 3, 4))
    out = tmp_path / "m"
    n = C.merge([tmp_path / "a", tmp_path / "b"], out, ng, copy_conf=True)
    assert n == 5
    lines = open(out / "eval.par").read().splitlines()
    # energy, then sigma breaks the 0.05 tie (b's 1e-5 before a's 5e-5)
    want = [(-0.20, 3e-4, b[1]), (-0.10, 2e-4, a[1]), (0.05, 1e-5, b[0]), (0.05, 5e-5, a[2]),
            (0.30, 1e-3, a[0])]
    assert lines == ["%d %.16g %.16g" % (k, e, s) for k, (e, s, _) in enumerate(want)]
    assert lines[0] == "0 -0.2 0.0003"
    psi = np.fromfile(out / "psi.par").reshape(5, ng)
    for k, (_, _, blk) in enumerate(want):
        assert np.array_equal(psi[k], blk)
    man = open(out / "merge_manifest.txt").read().splitlines()
    assert man[0] == "ngrid 24" and man[1] == "nstates 5"
    assert man[2].startswith("source 0 ") and man[2].endswith("states=3")
    assert man[4] == "merged_state 0 source=1 local_index=1 energy=-0.2 sigma=0.0003"
    assert open(out / "conf.par").read() == open(tmp_path / "a" / "conf.par").read()


def test_refusals(tmp_path):
    ng = 8
    make_run(tmp_path / "ok", [0.1, 0.2], [1e-4, 1e-4], ng, 3)
    # psi.dat of the wrong size
    make_run(tmp_path / "short", [0.1, 0.2], [1e-4, 1e-4], ng, 4)
    np.zeros(ng).tofile(tmp_path / "short" / "psi.dat")
    with pytest.raises(C.MergeError, match="bytes"):
        C.merge([tmp_path / "short"], tmp_path / "o1", ng)
    # indices not 0..n-1
    make_run(tmp_path / "idx", [0.1, 0.2], [1e-4, 1e-4], ng, 5)
    with open(tmp_path / "idx" / "eval.dat", "w") as f:
        f.write("1 0.1 1e-4\n0 0.2 1e-4\n")
    with pytest.raises(C.MergeError, match="indices"):
        C.merge([tmp_path / "idx"], tmp_path / "o2", ng)
    # different structures
    make_run(tmp_path / "other", [0.3], [1e-4], ng, 6, conf="1\nCd 0 0 0\n")
    with pytest.raises(C.MergeError, match="conf.par"):
        C.merge([tmp_path / "ok", tmp_path / "other"], tmp_path / "o3", ng)
    # grid of the run's input.par
    make_run(tmp_path / "grid", [0.1], [1e-4], ng, 7, nxyz=(2, 2, 3))
    with pytest.raises(C.MergeError, match="ngrid"):
        C.merge([tmp_path / "grid"], tmp_path / "o4", ng)
    # missing files, short lines
    os.makedirs(tmp_path / "empty")
    with pytest.raises(C.MergeError, match="missing"):
        C.merge([tmp_path / "empty"], tmp_path / "o5", ng)
    make_run(tmp_path / "cols", [0.1], [1e-4], ng, 8)
    with open(tmp_path / "cols" / "eval.dat", "w") as f:
        f.write("0 0.1\n")
    with pytest.raises(C.MergeError, match="columns"):
        C.merge([tmp_path / "cols"], tmp_path / "o6", ng)


def test_cli(tmp_path, capsys):
    ng = 12
    make_run(tmp_path / "r", [0.2, 0.1], [1e-4, 2e-4], ng, 9, nxyz=(2, 2, 3))
    inp = tmp_path / "r" / "input.par"
    assert C.main([str(tmp_path / "r"), "--grid-source", str(inp),
                   "--output-dir", str(tmp_path / "out")]) == 0
    assert open(tmp_path / "out" / "eval.par").read().splitlines()[0] == "0 0.1 0.0002"
    assert C.main([str(tmp_path / "r"), "--ngrid", "13", "--output-dir",
                   str(tmp_path / "bad")]) == 1
    assert "error:" in capsys.readouterr().err
    with pytest.raises(SystemExit):
        C.main([str(tmp_path / "r")])  # neither --grid-source nor --ngrid


def test_couplings_driver_rejects_bad_options(tmp_path):
    """bin/couplings refuses malformed options before touching a device,
    writing the reason to `error` as the filter driver does."""
    import subprocess
    from semi_empirical_pseudopotentials_amd._lib import COUPLINGS_BIN
    for args, msg in ((["--fermi-energy", "np.float64(0.1)"], "needs a number"),
                      (["--state-type", "both"], "adiabatic or diabatic"),
                      (["--frobnicate"], "unknown option")):
        r = subprocess.run([COUPLINGS_BIN, "--dir", str(tmp_path), *args], capture_output=True,
                           text=True, timeout=60)
        assert r.returncode == 1
        assert msg in open(tmp_path / "error").read()
    # a directory without a deck
    r = subprocess.run([COUPLINGS_BIN, "--dir", str(tmp_path)], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1


def test_filter_driver_rejects_bad_lanes(tmp_path):
    """bin/filter refuses a --lanes value outside 0..SEP_MAX_LANES before
    reading the deck or touching a device."""
    import subprocess
    from semi_empirical_pseudopotentials_amd._lib import FILTER_BIN
    for bad in ("9", "-1", "two"):
        r = subprocess.run([FILTER_BIN, "--dir", str(tmp_path), "--lanes", bad],
                           capture_output=True, text=True, timeout=60)
        assert r.returncode == 1
        assert "--lanes must be" in open(tmp_path / "error").read()


def test_filter_driver_rejects_unknown_options(tmp_path):
    """bin/filter refuses an unknown option (a misspelt --precision would
    otherwise run FP64 unnoticed) and a bad --refine-passes, before reading
    the deck or touching a device."""
    import subprocess
    from semi_empirical_pseudopotentials_amd._lib import FILTER_BIN
    for args, msg in ((["--precison", "fp32"], "unknown option or missing value: --precison"),
                      (["--refine-passes", "x"], "--refine-passes must be"),
                      (["--refine-passes", "-2"], "--refine-passes must be"),
                      (["--seed"], "unknown option or missing value: --seed")):
        r = subprocess.run([FILTER_BIN, "--dir", str(tmp_path), *args], capture_output=True,
                           text=True, timeout=60)
        assert r.returncode == 1, args
        assert msg in open(tmp_path / "error").read(), args
