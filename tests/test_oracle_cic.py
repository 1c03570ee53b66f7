"""Pins of the oracle's CIC deposit and CIC^-1 (main file:1077-1390).

Each check is chosen so that a plausible restatement mistake fails it: a
wrong node index or wrap (closed forms, brute force with an independent
tent-function formula), a dropped 1/dx^3 (mass conservation), a weight
transposed between axes (non-symmetric brute-force positions), and a
deposit/interpolation inconsistency (adjointness).
"""
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def tent_grid(pos, ng, xxmin=0.0):
    """rho/rho_bar from the tent kernel: weight of node n is
    prod_axis max(0, 1 - |x - x_n|/dx) over periodic images."""
    dx = 1.0 / ng
    nodes = xxmin + np.arange(ng) * dx
    grid = np.zeros((ng, ng, ng))
    m = np.float32(1.0 / len(pos))
    for p in pos.astype(np.float64):
        w = []
        for a in range(3):
            d = np.abs(p[a] - nodes)
            d = np.minimum(d % 1.0, 1.0 - (d % 1.0))
            w.append(np.maximum(0.0, 1.0 - d / dx))
        grid += float(m) * np.einsum("k,j,i->kji", w[2], w[1], w[0])
    return grid / dx ** 3


def test_closed_forms(orc):
    g = json.load(open(os.path.join(GOLD, "cic_closed_form.json")))
    ng = g["ng"]
    grid = orc.cic(np.array([g["on_node"]["pos"]], np.float32), ng)
    i, j, k = g["on_node"]["node"]
    assert grid[k, j, i] == g["on_node"]["value"]
    assert np.count_nonzero(grid) == 1
    grid = orc.cic(np.array([g["centre"]["pos"]], np.float32), ng)
    for i, j, k in g["centre"]["nodes"]:
        assert grid[k, j, i] == g["centre"]["value"]
    assert np.count_nonzero(grid) == 8


@pytest.mark.parametrize("ng,xxmin", [(8, 0.0), (8, -0.5), (16, -0.25)])
def test_brute_force_tent(orc, ng, xxmin):
    rng = np.random.default_rng(1)
    pos = rng.random((37, 3)).astype(np.float32)
    pos[0] = [0.999999, 0.0, 0.5]        # periodic wrap in x
    pos[1] = [0.3, 0.7, 0.11]            # asymmetric: catches axis swaps
    ref = tent_grid(pos, ng, xxmin)
    got = orc.cic(pos, ng, xxmin=xxmin)
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-9)


def test_mass_conservation_and_translation(orc):
    rng = np.random.default_rng(2)
    ng = 16
    pos = rng.random((500, 3)).astype(np.float32)
    g0 = orc.cic(pos, ng)
    # total mass is N * float32(1/N), the reference's real(4) mpf (:662)
    m = float(np.float32(1.0 / len(pos))) * len(pos)
    assert abs(g0.sum() / ng ** 3 - m) < 1e-12
    # a shift by whole cells permutes the grid (np.roll)
    sh = np.array([3, 5, 1]) / ng
    pos2 = ((pos.astype(np.float64) + sh) % 1.0).astype(np.float32)
    ok = np.all(np.abs(pos2.astype(np.float64) - ((pos + sh) % 1.0)) == 0, axis=1)
    g1 = orc.cic(pos2[ok], ng)
    g0b = orc.cic(pos[ok], ng)
    np.testing.assert_allclose(g1, np.roll(g0b, (1, 5, 3), axis=(0, 1, 2)),
                               rtol=1e-12, atol=1e-12)


def test_zone_origin_relabels_nodes(orc):
    rng = np.random.default_rng(3)
    ng = 8
    pos = rng.random((200, 3)).astype(np.float32)
    a = orc.cic(pos, ng, xxmin=0.0)
    b = orc.cic(pos, ng, xxmin=-0.5)  # the reference's single-zone origin
    np.testing.assert_allclose(b, np.roll(a, (4, 4, 4), axis=(0, 1, 2)),
                               rtol=1e-12)


def test_cicm1_uniform_lattice_and_adjoint(orc):
    ng = 8
    ax = (np.arange(ng) + 0.0) / ng
    lat = np.array(np.meshgrid(ax, ax, ax, indexing="ij")).reshape(3, -1).T
    lat = lat.astype(np.float32)
    grid = orc.cic(lat, ng)
    np.testing.assert_allclose(grid, 1.0, rtol=1e-12)
    rho = orc.cicm1(lat, grid)
    np.testing.assert_allclose(rho, 1.0, rtol=1e-12)
    # adjointness: sum_i m rho_i(g) == sum_n cic(m)_n g_n dx^3
    rng = np.random.default_rng(4)
    pos = rng.random((300, 3)).astype(np.float32)
    g = rng.random((ng, ng, ng))
    lhs = orc.cicm1(pos, g).sum() * float(np.float32(1.0 / len(pos)))
    rhs = (orc.cic(pos, ng) * g).sum() / ng ** 3
    assert abs(lhs - rhs) < 1e-12 * max(1.0, abs(lhs))
