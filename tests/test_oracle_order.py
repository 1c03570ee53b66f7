"""Pins of the oracle's quick_sort, tri_cic and cell bucketing
(main file:2477-2571, 998-1037, 1561-1604)."""
import json
import os

import numpy as np

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_quick_sort_hand_trace(orc):
    g = json.load(open(os.path.join(GOLD, "quick_sort_trace.json")))
    lst, order = orc.quick_sort(g["input"])
    assert list(lst) == g["sorted"]
    # the published method is not stable: its tie order is the hand trace
    assert list(order) == g["order"]
    assert list(order) != g["stable_order"]


def test_quick_sort_random_is_a_sorting_permutation(orc):
    rng = np.random.default_rng(5)
    for n in (0, 1, 2, 6, 7, 8, 100, 5000):
        v = rng.integers(0, 20, size=n).astype(np.float64)  # many ties
        lst, order = orc.quick_sort(v)
        assert np.all(np.diff(lst) >= 0)
        assert sorted(order) == list(range(1, n + 1))
        np.testing.assert_array_equal(v[order - 1], lst)


def test_tri_cic_count_and_order(orc):
    rng = np.random.default_rng(6)
    rho = np.exp(rng.normal(2.0, 2.5, size=4000))
    pdens, isort, nnew = orc.tri_cic(rho, 100.0)
    assert nnew == int((rho > 100.0).sum()) + 1
    np.testing.assert_array_equal(pdens, np.sort(1.0 / rho))
    assert np.all(np.diff(rho[isort - 1]) <= 0)


def test_cell_sort_brute_force(orc):
    rng = np.random.default_rng(7)
    nx, bs = 16, 8
    pos = rng.random((3000, 3)).astype(np.float32)
    pos[:5] = [[0.0, 0.0, 0.0], [0.9999999, 0.5, 0.5], [0.5, 0.0625, 0.5],
               [0.25, 0.25, 0.25], [0.4375, 0.9375, 0.0]]
    cs, perm = orc.cell_sort(pos, nx, bs)
    cells = np.floor(pos.astype(np.float64) * nx).astype(int) % nx
    key = np.array([orc.cell_key(nx, bs, *c) for c in cells])
    hist = np.bincount(key, minlength=nx ** 3)
    np.testing.assert_array_equal(cs[1:] - cs[:-1], hist)
    assert cs[0] == 0 and cs[-1] == len(pos)
    for c in range(0, nx ** 3, 37):
        np.testing.assert_array_equal(perm[cs[c]:cs[c + 1]],
                                      np.where(key == c)[0])


def test_cell_key_is_a_blocked_bijection(orc):
    nx, bs = 16, 8
    keys = {orc.cell_key(nx, bs, i, j, k)
            for i in range(nx) for j in range(nx) for k in range(nx)}
    assert keys == set(range(nx ** 3))
    # the 8^3 cells of block 0 are the first 512 keys
    assert orc.cell_key(nx, bs, 7, 7, 7) == 511
    assert orc.cell_key(nx, bs, 8, 0, 0) == 512
