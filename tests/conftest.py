import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def tk():
    import __graft_entry__
    return __graft_entry__.package()


def to_sym(tk, zt):
    """oracle charge-labelled tensor -> engine SymTensor (space labels:
    codomain legs q, domain legs -q)."""
    N = zt.N
    labels = [zt.q[0], zt.q[1], (-zt.q[2]) % N, (-zt.q[3]) % N]
    return tk.SymTensor.from_dense(zt.data, labels, N)


def rel(a, b):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    return float(np.max(np.abs(a / b - 1)))
