import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")
    config.addinivalue_line("markers", "slow: CPU test that takes tens of seconds")


@pytest.fixture(scope="session")
def oracle():
    from oracle.pyoracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def ref_oracle():
    from oracle.pyoracle import REF_SO, Oracle
    if not os.path.exists(REF_SO):
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    return Oracle(ref=True)


def random_state(n, seed, dtype=np.complex128):
    rs = np.random.default_rng(seed)
    v = rs.normal(size=1 << n) + 1j * rs.normal(size=1 << n)
    v /= np.linalg.norm(v)
    return v.astype(dtype)


def rel_err(a, b, floor=1.0):
    """max |a-b| / max(|b|, floor): relative error with an absolute floor for
    entries near zero (the tolerance contract is 1e-12 relative, complex128)."""
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))) if b.size else 0.0, floor))
