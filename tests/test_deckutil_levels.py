"""The one-to-one level comparison of tests/deckutil.py on planted cases:
an exact match, a lost multiplet member, two levels merged into one, and an
extra level inside a multiplet must be told apart (nearest-value matching
passes all four)."""
import numpy as np
import pytest

from tests.deckutil import compare_levels

TOL = 1e-3 / 27.211386245988
REF = np.array([-0.2, -0.15, -0.15, -0.15, -0.1, -0.05, -0.05])
SIG = np.full(REF.size, 1e-5)


def test_match_passes_and_counts():
    got = REF + 0.3 * TOL * np.sin(np.arange(REF.size))
    nc, nl, d = compare_levels(REF, SIG, got, SIG, -1.0, 1.0, TOL)
    assert (nc, nl) == (4, 7) and d < TOL


def test_lost_level_fails():
    got = np.delete(REF, 2)
    with pytest.raises(AssertionError):
        compare_levels(REF, SIG, got, SIG[:-1], -1.0, 1.0, TOL)


def test_merged_levels_fail():
    got = np.delete(REF, 6)  # the pair at -0.05 became one level
    with pytest.raises(AssertionError):
        compare_levels(REF, SIG, got, SIG[:-1], -1.0, 1.0, TOL)


def test_extra_level_fails_but_loose_sigma_excludes_noise():
    got = np.append(REF, -0.15 + 0.2 * TOL)
    with pytest.raises(AssertionError):
        compare_levels(REF, SIG, got, np.full(got.size, 1e-5), -1.0, 1.0, TOL)
    # the same extra value with a noise-level variance is not a level
    sg = np.append(SIG, 0.5)
    nc, nl, d = compare_levels(REF, SIG, got, sg, -1.0, 1.0, TOL)
    assert nl == 7


def test_unconverged_reference_clusters_are_skipped():
    s = SIG.copy()
    s[4] = 5e-3  # -0.1 resolved only loosely in the reference
    got = np.delete(REF, 4)
    nc, nl, d = compare_levels(REF, s, got, np.full(got.size, 1e-5), -1.0, 1.0, TOL)
    assert (nc, nl) == (3, 6)
