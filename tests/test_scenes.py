"""Scene generators: sizes follow the workload arithmetic (two n x n sheets:
V = 2 n^2, E = 2 (2 n (n-1) + (n-1)^2), F = 4 (n-1)^2) and are seeded."""
import numpy as np
import pytest


@pytest.mark.parametrize("n", [2, 5, 64])
def test_two_sheets_sizes(scenes, n):
    s = scenes.two_sheets(n, seed=0)
    assert s.sizes == (2 * n * n, 2 * (2 * n * (n - 1) + (n - 1) ** 2), 4 * (n - 1) ** 2)


def test_config_sizes():
    # 2 x 64^2 and 8 x 512^2 sheets
    V = lambda n, k: k * n * n
    E = lambda n, k: k * (2 * n * (n - 1) + (n - 1) ** 2)
    F = lambda n, k: k * 2 * (n - 1) ** 2
    assert (V(64, 2), E(64, 2), F(64, 2)) == (8192, 24066, 15876)
    assert (V(512, 8), E(512, 8), F(512, 8)) == (2097152, 6275080, 4177936)


def test_stacked_sheets_small(scenes):
    s = scenes.stacked_sheets(9, 4, seed=3)
    n = 9
    assert s.sizes == (4 * n * n, 4 * (2 * n * (n - 1) + (n - 1) ** 2), 8 * (n - 1) ** 2)
    z = s.V[:, 2].reshape(4, -1)
    assert np.all(np.abs(z - (0.6e-3 * np.arange(4))[:, None]) <= 0.2e-3 + 1e-18)
    d = s.V1 - s.V
    assert np.allclose(d[:, 2], -0.8e-3)


def test_seeded(scenes):
    a, b = scenes.stacked_sheets(8, 3, seed=5), scenes.stacked_sheets(8, 3, seed=5)
    assert np.array_equal(a.V, b.V) and np.array_equal(a.V1, b.V1)
    c = scenes.stacked_sheets(8, 3, seed=6)
    assert not np.array_equal(a.V, c.V)


def test_edges_unique_sorted(scenes):
    s = scenes.sphere_chain(2, 2, seed=0)
    E = s.E
    assert np.all(E[:, 0] < E[:, 1])
    key = E[:, 0].astype(np.int64) << 32 | E[:, 1]
    assert np.all(np.diff(key) > 0)
    # closed surfaces: every edge has two faces (Euler: V - E + F = 2 per sphere)
    assert len(s.V) - len(E) + len(s.F) == 2 * 2
