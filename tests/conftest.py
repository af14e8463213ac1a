import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (run with -m gpu)")


@pytest.fixture(scope="session")
def ctx():
    """One pray context for all GPU tests (fails loudly without a GPU)."""
    import nostrum_pray_amd as P
    c = P.Context(0)
    yield c
    c.close()


def say(msg):
    """Progress line naming the case, flushed before GPU work starts."""
    print(msg, flush=True)


def first_diff(a, b):
    d = np.flatnonzero(a != b)
    return None if d.size == 0 else (int(d[0]), a[d[0]], b[d[0]], int(d.size))
