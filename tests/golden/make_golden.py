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
  powerlaw_crossing.npz  O(k) an exact power law in r: the log-log crossing
                         radius and mass in closed form.
  com_edge.npz           a sphere of symmetric pairs straddling the periodic
                         edge: its centre of mass is the centre, exactly.
  greedy_pair.npz        two overlapping halos: the second one's count with
                         the first one's sphere excluded is known.

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


V0 = 2.0 / 3.0 * 2.0 * np.pi  # (2/3)(2 pi): sphere volume / r^3


def fill_background(rng, nx, n, centres, clear):
    """Uniform particles more than `clear` cells from every centre."""
    bg = []
    while len(bg) < n:
        x = rng.uniform(0, nx, size=3)
        ok = True
        for c in centres:
            d = np.abs(x - c)
            d = np.minimum(d, nx - d)
            ok = ok and np.sqrt((d ** 2).sum()) > clear
        if ok:
            bg.append(x)
    return np.array(bg).reshape(-1, 3)


def dyadic(v, bits=12):
    """Coordinates rounded to multiples of 2^-bits cells: exact in float32
    box units for nx = 16 (x = X/16 with X a multiple of 2^-12)."""
    return np.round(np.asarray(v) * 2.0 ** bits) / 2.0 ** bits


def unit(rng, hemi=0.0):
    """Isotropic unit vector; hemi > 0 keeps u_x >= hemi."""
    while True:
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        if v[0] >= hemi:
            return v


def powerlaw(nx=16, delta=200.0, kcross=20.5, kmax=40, seed=5):
    """Particle k >= 10 at r_k = k*u, u = 1/(V0*A), so that the growth's
    O(k) = k/(V0 r_k^3) = A r_k^-2 is an exact power law in r (particles
    2..9 form a cube-corner knot 0.017 cells from the centre particle, far
    above Delta, which makes the centre the clear maximum of S).  Log-log
    interpolation between the two particles around the crossing is then
    exact: r* = (A/Delta)^(1/2), M* = Delta (4 pi/3) r*^3, where
    A^3 = Delta kcross^2 / V0^2 puts O(kcross) = Delta.  Linear
    interpolation in r gives a radius 3e-4 relative larger."""
    rng = np.random.default_rng(seed)
    A = (delta * kcross ** 2 / V0 ** 2) ** (1.0 / 3.0)
    u = 1.0 / (V0 * A)
    c = np.array([8.5, 8.5, 8.5])
    pts = ([c.copy()] + knot(c, [0.01]) +
           [c + k * u * unit(rng) for k in range(10, kmax + 1)])
    halo = np.array(pts)
    bg = fill_background(rng, nx, nx ** 3 - len(halo), [c],
                         kmax * u + 1.0)
    pos = (np.vstack([halo, bg]) / nx).astype(np.float32)
    pos[pos >= 1.0] = 0.0
    lab = np.full(nx ** 3, -1, np.int32)
    lab[:len(halo)] = 0
    kstar = int(np.ceil(kcross))
    rstar = np.sqrt(A / delta)
    return dict(pos=pos, labels=lab, nx=np.int32(nx), delta=np.float64(delta),
                centre=np.int64(0), count=np.int64(kstar - 1),
                A=np.float64(A), u=np.float64(u), rint=np.float64(rstar),
                mint=np.float64(delta * 4.0 / 3.0 * np.pi * rstar ** 3))


def com_edge(nx=16, delta=200.0, npair=4, seed=6):
    """A halo centred 1/16 cell from the x = 0 face, so that its sphere
    straddles the periodic edge: the centre plus symmetric pairs c +- v,
    every coordinate a multiple of 2^-12 cells (exact in float32).  After
    the knot (k = 2..9), pairs 1..npair-1 sit where O(9+2i) = 2*Delta,
    pair npair where O = Delta/2: the sphere holds the centre, the knot and
    npair-1 whole pairs, and their mean is the centre exactly.
    com_pos = c / nx, with c = (1/16, 8.5, 8.5)."""
    rng = np.random.default_rng(seed)
    c = np.array([0.0625, 8.5, 8.5])
    # a cube-corner knot (4 more symmetric pairs) makes the centre the
    # clear maximum of S
    pts = [c.copy()] + knot(c, [dyadic(0.01)])
    for i in range(1, npair + 1):
        k = 9 + 2 * i
        a = 2.0 if i < npair else 0.5
        # O(k) = k/(V0 r^3) = a*Delta at the outer member; the inner member
        # (k - 1) is at the same radius
        r = (k / (V0 * a * delta)) ** (1.0 / 3.0)
        v = dyadic(r * unit(rng))
        pts.append(c + v)
        pts.append(c - v)
    halo = np.array(pts) % nx
    bg = dyadic(fill_background(rng, nx, nx ** 3 - len(halo), [c], 3.0))
    pos = (np.vstack([halo, bg]) / nx).astype(np.float32)
    pos[pos >= 1.0] = 0.0
    assert np.all(pos.astype(np.float64) * nx == np.vstack([halo, bg]) % nx)
    lab = np.full(nx ** 3, -1, np.int32)
    lab[:len(halo)] = 0
    return dict(pos=pos, labels=lab, nx=np.int32(nx), delta=np.float64(delta),
                centre=np.int64(0), count=np.int64(9 + 2 * (npair - 1)),
                com_pos=(c / nx).astype(np.float64))


def knot(c, radii):
    """Cube-corner shells around c (8 particles per radius)."""
    s = np.array([[i, j, k] for i in (-1, 1) for j in (-1, 1)
                  for k in (-1, 1)], float)
    return [c + r * v for r in radii for v in s]


def greedy_pair(nx=16, delta=200.0, seed=7):
    """Two halos for the greedy pass (main file:1617-2166), B 0.45 cells
    from A along +x.  A (denser: a 24-particle knot) is grown first; its
    sphere holds A's centre, knot and shells k = 26..31 at O = 2*Delta,
    and its crossing particle (k = 32, O = Delta/2) lies on the -x side.
    A's shells include one particle on the +x axis, 0.21 cells from B's
    centre.  B (16-particle knot, shells k = 18..27 at O = 2*Delta on the
    +x side, crossing k = 28 at Delta/2) is grown second from the alive
    particles only: with A's sphere excluded, B's count is its design, 27.
    Counting A's particles too would put the one at 0.21 cells inside B's
    radius and give B 28 particles."""
    rng = np.random.default_rng(seed)
    cA = np.array([8.5, 8.5, 8.5])
    cB = cA + np.array([0.45, 0.0, 0.0])

    def shell_r(k, a):
        return (k / (V0 * a * delta)) ** (1.0 / 3.0)

    a_pts = [cA.copy()] + knot(cA, [0.02, 0.04, 0.06])        # k = 1..25
    ka = len(a_pts)
    a_sh = []
    for k in range(ka + 1, 32):                                # k = 26..31
        d = np.array([1.0, 0.0, 0.0]) if k == 31 else unit(rng)
        a_sh.append(cA + shell_r(k, 2.0) * d)
    a_cross = cA + shell_r(32, 0.5) * np.array([-1.0, 0.0, 0.0])
    b_pts = [cB.copy()] + knot(cB, [0.03, 0.06])              # k = 1..17
    kb = len(b_pts)
    b_sh = [cB + shell_r(k, 2.0) * unit(rng, hemi=0.5)
            for k in range(kb + 1, 28)]                        # k = 18..27
    b_cross = cB + shell_r(28, 0.5) * np.array([1.0, 0.0, 0.0])
    A = np.array(a_pts + a_sh)
    B = np.array(b_pts + b_sh)
    halo = np.vstack([A, [a_cross], B, [b_cross]])
    bg = fill_background(rng, nx, nx ** 3 - len(halo), [cA, cB], 3.0)
    pos = (np.vstack([halo, bg]) / nx).astype(np.float32)
    pos[pos >= 1.0] = 0.0
    nA, nB = len(A), len(B)
    ia = np.arange(nA)                       # A's sphere
    ib = np.arange(nA + 1, nA + 1 + nB)      # B's sphere
    # the design, checked on the stored float32 positions
    xs = pos.astype(np.float64) * nx
    ra = np.sort(np.sqrt(((xs[:len(halo)] - xs[0]) ** 2).sum(1)))
    assert ra[nA - 1] < np.sqrt(((xs[nA] - xs[0]) ** 2).sum())
    rb_all = np.sqrt(((xs[:len(halo)] - xs[nA + 1]) ** 2).sum(1))
    rb = np.sqrt(((xs[ib] - xs[nA + 1]) ** 2).sum(1))
    assert rb.max() < rb_all[nA + 1 + nB]               # B's crossing particle
    assert np.sqrt(((xs[30] - xs[nA + 1]) ** 2).sum()) < rb.max()  # A's +x
    assert np.all(np.sqrt(((xs[ib] - xs[0]) ** 2).sum(1)) > ra[nA - 1])
    return dict(pos=pos, nx=np.int32(nx), delta=np.float64(delta),
                centre_a=np.int64(0), count_a=np.int64(nA),
                centre_b=np.int64(nA + 1), count_b=np.int64(nB),
                sphere_a=ia.astype(np.int64), sphere_b=ib.astype(np.int64),
                radius_a=np.float64(np.sort(np.sqrt(((xs[ia] - xs[0]) ** 2)
                                                    .sum(1)))[-1]),
                radius_b=np.float64(rb.max()))


def main():
    with open(os.path.join(HERE, "quick_sort_trace.json"), "w") as f:
        json.dump(quick_sort_trace(), f, indent=1)
    with open(os.path.join(HERE, "cic_closed_form.json"), "w") as f:
        json.dump(cic_closed_form(), f, indent=1)
    np.savez_compressed(os.path.join(HERE, "shells_crossing.npz"), **shells())
    np.savez_compressed(os.path.join(HERE, "shells_floor.npz"),
                        **shells(floor_case=True))
    np.savez_compressed(os.path.join(HERE, "powerlaw_crossing.npz"),
                        **powerlaw())
    np.savez_compressed(os.path.join(HERE, "com_edge.npz"), **com_edge())
    np.savez_compressed(os.path.join(HERE, "greedy_pair.npz"),
                        **greedy_pair())


if __name__ == "__main__":
    main()
