"""The C ABI library: it loads without a GPU, exports every function that
include/*.h declares, and the driver fails the reference's way on bad input."""
import os
import re
import subprocess

import pytest

from tests.conftest import ROOT


def _declared():
    names = set()
    for h in ("sep.h", "sep_deck.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(sep_\w+)\s*\(", txt, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_symbols_exported(sep):
    import ctypes
    from semi_empirical_pseudopotentials_amd._lib import LIB_PATH, SIGNATURES
    decl = _declared()
    assert len(decl) >= 25
    L = ctypes.CDLL(LIB_PATH)
    missing = [n for n in decl if not hasattr(L, n)]
    assert not missing, missing
    assert decl <= set(SIGNATURES), decl - set(SIGNATURES)


def test_open_without_gpu_fails_cleanly(sep):
    """No device here: sep_open returns SEP_ERR_HIP instead of crashing."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(sep.SepError):
        sep.Context(16, 16, 16, 1.0, 1.0, 1.0)


def test_unsupported_length(sep):
    assert sep.fft_length_supported(64) and sep.fft_length_supported(36)
    assert not sep.fft_length_supported(17)


def test_driver_missing_input(sep, tmp_path):
    """A deck without conf.par writes `error` and exits 1 (filter/nerror.c:9-16)."""
    from semi_empirical_pseudopotentials_amd._lib import FILTER_BIN
    (tmp_path / "input.par").write_text("16 16 16 2 1 8 -0.3 -0.1 1\n")
    r = subprocess.run([FILTER_BIN, "--dir", str(tmp_path)], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1
    assert "conf.par" in (tmp_path / "error").read_text()


def test_driver_unknown_atom(sep, tmp_path):
    from semi_empirical_pseudopotentials_amd._lib import FILTER_BIN
    (tmp_path / "input.par").write_text("16 16 16 2 1 8 -0.3 -0.1 1\n")
    (tmp_path / "conf.par").write_text("1\nXx 0 0 0\n")
    r = subprocess.run([FILTER_BIN, "--dir", str(tmp_path)], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1
    assert "not in current list" in (tmp_path / "error").read_text()
