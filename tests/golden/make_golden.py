"""Writes the golden fixtures of tests/golden/ from closed forms.

Nothing here calls the oracle or the HIP library: every expected value is
fixed by construction or by a hand trace, so the fixtures can pin both.

  quick_sort_trace.json  input with ties and the permutation produced by the
                         published quicksort (Brainerd, Goldberg & Adams 1990,
                         pp. 149-150, with A. Miller's index array) -- traced
                         by hand, step by step, in the comments below.
  cic_closed_form.json   particles on a node and at a cell centre.
  shells_*.npz           a halo whose particles sit on designed radii so that
                         the first k >= 10 with O(k) <= Delta is known.

Run: python tests/golden/make_golden.py
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def quick_sort_trace():
    # list = [3,1,3,2,3,1,2,3], n = 8 (1-based positions), order = 1..8
    # qs(1,8): 8 >= 1+6 -> partition, pivot = list((1+8)/2) = list(4) = 2
    #   i=1 (3>=2), j=7 (2<=2): swap -> list [2,1,3,2,3,1,3,3] order [7,2,3,4,5,6,1,8]
    #   i=3 (3>=2), j=6 (1<=2): swap -> list [2,1,1,2,3,3,3,3] order [7,2,6,4,5,3,1,8]
    #   i=4 (2>=2), j=4 (2<=2): i==j -> i=5, exit
    #   qs(1,4) and qs(5,8)
    # qs(1,4): interchange sort of [2,1,1,2] / [7,2,6,4]
    #   (1,2): 2>1 swap -> [1,2,1,2] [2,7,6,4]; (2,3): 2>1 swap -> [1,1,2,2]
    #   [2,6,7,4]; no other swap
    # qs(5,8): [3,3,3,3] unchanged, order [5,3,1,8]
    return dict(input=[3, 1, 3, 2, 3, 1, 2, 3],
                sorted=[1, 1, 2, 2, 3, 3, 3, 3],
                order=[2, 6, 7, 4, 5, 3, 1, 8],
                stable_order=[2, 6, 4, 7, 1, 3, 5, 8])


def cic_closed_form():
    # ng = 4, one particle of mass 1/N with N = 1 (mpf = 1/nparttotal):
    # on node (1, 2, 3) -> the whole mass on that node: 1 / dx^3 = 64;
    # at a cell centre -> 8 nodes with 64/8 = 8 each.
    return dict(
        ng=4,
        on_node=dict(pos=[0.25, 0.5, 0.75], node=[1, 2, 3], value=64.0),
        centre=dict(pos=[0.125, 0.625, 0.875],
                    nodes=[[i, j, k] for i in (0, 1) for j in (2, 3)
                           for k in (3, 0)],
                    value=8.0),
    )


def shells(nx=16, delta=200.0, kstar=40, kmax=120, knot=0.05, seed=3,
           floor_case=False):
    """A halo with a designed crossing.  Mass per particle is 1 in units of
    rho_bar*cell^3, i.e. the box holds nx^3 particles: the halo plus a
    uniform background kept more than 3 cells from the centre."""
    rng = np.random.default_rng(seed)
    c = np.array([8.5, 8.5, 8.5])  # cells; 8.5/16 is exact in float32
    pts = [c.copy()]
    if floor_case:
        # 8 particles at ~0.33 cells: O(9) = 9/(4/3 pi 0.33^3) < Delta, so
        # only the 10-particle floor keeps the sphere growing to k = 10
        for s in np.array([[i, j, k] for i in (-1, 1) for j in (-1, 1)
                           for k in (-1, 1)], float):
            pts.append(c + 0.19 * s)
        radii = [0.6 + 0.05 * i for i in range(kmax - 9)]
    else:
        for s in np.array([[i, j, k] for i in (-1, 1) for j in (-1, 1)
                           for k in (-1, 1)], float):
            pts.append(c + knot * s)
        radii = []
        for k in range(10, kmax + 1):
            a = 2.0 if k < kstar else (0.5 if k == kstar else 0.3)
            radii.append((3.0 * k / (4.0 * np.pi * a * delta)) ** (1.0 / 3.0))
        radii = np.maximum.accumulate(np.array(radii) * (1 + 1e-3 * np.arange(len(radii)) / len(radii)))
    for r in radii:
        v = rng.normal(size=3)
        pts.append(c + r * v / np.linalg.norm(v))
    halo = np.array(pts)
    n = nx ** 3
    bg = []
    while len(bg) < n - len(halo):
        x = rng.uniform(0, nx, size=3)
        d = np.abs(x - c)
        d = np.minimum(d, nx - d)
        if np.sqrt((d ** 2).sum()) > 3.0:
            bg.append(x)
    allp = np.vstack([halo, np.array(bg)])
    pos = (allp / nx).astype(np.float32)
    pos[pos >= 1.0] = 0.0
    lab = np.full(n, -1, np.int32)
    lab[:len(halo)] = 0
    # exact distances of the stored float32 positions from the centre
    xs = pos.astype(np.float64) * nx
    d = np.abs(xs[:len(halo)] - xs[0])
    d = np.minimum(d, nx - d)
    r = np.sqrt((d ** 2).sum(axis=1))
    rs = np.sort(r)
    k = np.arange(1, len(rs) + 1)
    od = np.full(len(rs), np.inf)
    od[1:] = k[1:] / (4.0 / 3.0 * np.pi * rs[1:] ** 3)
    if floor_case:
        exp_kstar = 10
    else:
        exp_kstar = kstar
    assert od[exp_kstar - 1] <= delta and np.all(od[9:exp_kstar - 1] > delta)
    prof_r = rs[exp_kstar - 2] * np.exp(
        np.log(0.01) + np.arange(32) * (np.log(2.0) - np.log(0.01)) / 31.0)
    prof_r[-1] = rs[exp_kstar - 2] * 2.0  # the last radius is R * rmax exactly
    prof = np.searchsorted(rs, prof_r, side="left")
    return dict(pos=pos, labels=lab, nx=np.int32(nx), delta=np.float64(delta),
                centre=np.int64(0), count=np.int64(exp_kstar - 1),
                radius=np.float64(rs[exp_kstar - 2]), profile=prof.astype(np.int64))


def main():
    with open(os.path.join(HERE, "quick_sort_trace.json"), "w") as f:
        json.dump(quick_sort_trace(), f, indent=1)
    with open(os.path.join(HERE, "cic_closed_form.json"), "w") as f:
        json.dump(cic_closed_form(), f, indent=1)
    np.savez_compressed(os.path.join(HERE, "shells_crossing.npz"), **shells())
    np.savez_compressed(os.path.join(HERE, "shells_floor.npz"),
                        **shells(floor_case=True))


if __name__ == "__main__":
    main()
