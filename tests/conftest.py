import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "ipc-toolkit_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and libipcg.so")


@pytest.fixture(scope="session")
def O():
    """The CPU oracle (test infrastructure; built on first use)."""
    import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def capi():
    from ipctk_amd import capi as c
    return c


@pytest.fixture(scope="session")
def scenes():
    from ipctk_amd import scenes as s
    return s


@pytest.fixture()
def ctx(capi):
    """A device context on GPU 0 (GPU tests only)."""
    c = capi.Context(0)
    yield c
    c.close()


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b.ravel())
    num = np.linalg.norm((a - b).ravel())
    return num / den if den > 0 else num


def sparse_rel_err(A, B):
    D = (A - B).tocoo()
    nb = np.sqrt((B.multiply(B)).sum())
    nd = np.sqrt((D.multiply(D)).sum())
    return nd / nb if nb > 0 else nd


def pair_set(a):
    return set(map(tuple, np.asarray(a).reshape(-1, 2).tolist()))


def unordered_set(a):
    a = np.asarray(a).reshape(-1, 2)
    return set(map(tuple, np.sort(a, axis=1).tolist()))
