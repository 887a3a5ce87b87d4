import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP) device")
    config.addinivalue_line("markers", "slow: takes more than ~20 s on the CPU")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def sep():
    import semi_empirical_pseudopotentials_amd as S
    S.lib()
    return S


@pytest.fixture(scope="session")
def small_deck(tmp_path_factory):
    """1.5 nm CdSe dot on a 24^3 grid, short filter (fast on the CPU oracle)."""
    from semi_empirical_pseudopotentials_amd import decks
    d = tmp_path_factory.mktemp("deck_small")
    syms, xyz = decks.cdse_dot(1.5)
    decks.write_deck(str(d), syms, xyz, 24, 24, 24, ms=4, ns=3, nc=96, elmin=-0.35,
                     elmax=-0.15, seed=-1001)
    return str(d)
