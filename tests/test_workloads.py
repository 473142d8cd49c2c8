"""Synthetic molecule generators: geometry rules of the benchmark configs."""
import math

import numpy as np
import pytest

from molds_amd import workloads as W


def _dists(x):
    d = np.sqrt(((x[:, None, :] - x[None, :, :]) ** 2).sum(-1))
    np.fill_diagonal(d, np.inf)
    return d


def test_rng_is_deterministic_and_uniform():
    a = [W.Rng(42).uniform() for _ in range(3)]
    b = [W.Rng(42).uniform() for _ in range(3)]
    assert a == b
    r = W.Rng(7)
    u = np.array([r.uniform() for _ in range(20000)])
    assert abs(u.mean() - 0.5) < 0.01 and u.min() >= 0 and u.max() < 1
    g = np.array([r.normal() for _ in range(20000)])
    assert abs(g.mean()) < 0.03 and abs(g.std() - 1) < 0.03


@pytest.mark.parametrize("n", [1, 2, 5, 10, 40])
def test_alkane_geometry(n):
    Z, x = W.alkane(n, jitter=0.0)
    assert list(Z).count(6) == n and list(Z).count(1) == 2 * n + 2
    d = _dists(x)
    C = np.where(Z == 6)[0]
    H = np.where(Z == 1)[0]
    for h in H:   # every H sits 1.09 A from exactly one carbon
        assert abs(np.min(d[h, C]) - 1.09) < 1e-9
    for i in range(n - 1):
        assert abs(d[C[i], C[i + 1]] - 1.54) < 1e-9
    for c in C:   # tetrahedral carbons: four neighbours, 109.47 deg angles
        nb = np.argsort(d[c])[:4]
        v = x[nb] - x[c]
        v /= np.linalg.norm(v, axis=1)[:, None]
        ang = np.degrees(np.arccos(np.clip(v @ v.T, -1, 1)))[np.triu_indices(4, 1)]
        assert np.allclose(ang, W.TETRA, atol=1e-6)


def test_water_cluster_geometry():
    Z, x = W.water_cluster(64, seed=1)
    assert len(Z) == 192 and list(Z[:3]) == [8, 1, 1]
    O = x[Z == 8]
    dOO = _dists(O)
    assert dOO.min() >= 2.6
    L = (64 * 29.9) ** (1 / 3)
    assert O.min() >= 0 and O.max() < L
    for i in range(64):
        o, h1, h2 = x[3 * i], x[3 * i + 1], x[3 * i + 2]
        assert abs(np.linalg.norm(h1 - o) - 0.96) < 1e-12 and abs(np.linalg.norm(h2 - o) - 0.96) < 1e-12
        c = np.dot(h1 - o, h2 - o) / 0.96 ** 2
        assert abs(math.degrees(math.acos(c)) - 104.5) < 1e-9


def test_batch_mix_alternates_and_covers_sizes():
    items = list(W.batch_mix(40))
    assert [k for k, *_ in items[:4]] == ["alkane", "water", "alkane", "water"]
    for k, s, Z, x in items:
        if k == "alkane":
            assert 1 <= s <= 40 and len(Z) == 3 * s + 2
        else:
            assert 2 <= s <= 64 and len(Z) == 3 * s


def test_poisson_requests():
    reqs = list(W.poisson_requests(100.0, 200))
    t = np.array([r[0] for r in reqs])
    assert np.all(np.diff(t) > 0)
    rate = len(t) / t[-1]
    assert 70 < rate < 140
    for t0, k, s, Z, x in reqs:
        assert 4 <= len(Z) <= 602
