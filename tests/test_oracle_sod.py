"""Pins of the oracle's SOD (main file:1617-2166) in both modes.

Designed-radius halos fix the first crossing, the 10-particle floor and the
last-above rule of the second threshold by construction; the centre search
is checked against an independent numpy brute force; an isolated NFW clump
checks R_200 against its analytic value; computecom and the Delta
conversions against closed forms.
"""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def load(name):
    z = np.load(os.path.join(GOLD, name))
    return {k: z[k] for k in z.files}


def fof_run(orc, z, deltas=None, **kw):
    pos, nx = z["pos"], int(z["nx"])
    cs, perm = orc.cell_sort(pos, nx, 8)
    p = orc.fof_params(nx, len(pos), deltas=deltas or (float(z["delta"]),),
                       **kw)
    nmem = np.array([(z["labels"] == 0).sum()])
    return orc.sod_fof(pos, cs, perm, 8, np.array([int(z["centre"])]), nmem,
                       p)


@pytest.mark.parametrize("name", ["shells_crossing.npz", "shells_floor.npz"])
def test_fof_designed_crossing(orc, name):
    z = load(name)
    r = fof_run(orc, z)
    assert r["centre"][0] == z["centre"] and r["robust_centre"][0] == 1
    assert r["count"][0, 0] == z["count"]
    assert r["radius"][0, 0] == z["radius"]
    np.testing.assert_array_equal(r["profile"][0], z["profile"])
    if name == "shells_crossing.npz":
        # O(k*-1) > Delta >= O(k*): the crossing lies between r_{k*-1}, r_{k*}
        assert r["rint"][0, 0] >= r["radius"][0, 0]
        assert r["overdens"][0, 0] > float(z["delta"])
    else:
        # the 10-particle floor grew past the crossing: the log-log rule
        # extrapolates inside r_{k*-1} (haloanalyzer's "Will extrapolate")
        assert r["rint"][0, 0] < r["radius"][0, 0]


def test_second_threshold_last_above(orc):
    z = load("shells_crossing.npz")
    # knot (k = 2..9) is far above 800, the shells (k = 10..39) at 400
    assert fof_run(orc, z, delta2=800.0)["count2"][0] == 9
    assert fof_run(orc, z, delta2=300.0)["count2"][0] == int(z["count"])


def greedy_on(orc, z, **kw):
    pos, nx = z["pos"], int(z["nx"])
    grid = orc.cic(pos, 2 * nx, xxmin=-0.5)
    rho = orc.cicm1(pos, grid, xxmin=-0.5)
    _, isort, nnew = orc.tri_cic(rho)
    ids = np.arange(len(pos), dtype=np.int64)
    p = orc.greedy_params(nx, **kw)
    return orc.sod_greedy(pos, ids, isort, nnew, p)


def test_greedy_designed_crossing(orc):
    z = load("shells_crossing.npz")
    r = greedy_on(orc, z, mmin=20.0, seuil2=800.0)
    assert r["status"] == orc.OK
    h = r["halos"]
    k = int(np.where(h["centre"] == z["centre"])[0][0])
    assert h["masse"][k] == z["count"] and h["rayon"][k] == z["radius"]
    assert h["masse2"][k] == 9 and h["accepted"][k] == 1
    # the last profile radius is R itself (rmax = 1), so its count is the
    # particles strictly inside R: masse - 1 for distinct radii
    assert r["rad_bins"][k][-1] == h["rayon"][k]
    assert r["mass_bins"][k][-1] == h["masse"][k] - 1
    # the sphere's particles are excluded afterwards, accepted or not
    assert r["alive"][:int(z["count"])].sum() == 0


def test_greedy_abort_paths(orc):
    z = load("shells_crossing.npz")
    # Nsel overflow returns at once (main file:1681-1693)
    assert greedy_on(orc, z, mmin=20.0, nsel=50)["status"] == orc.E_CAPACITY
    # an sphere that cannot close inside ncut (main file:2072-2089)
    big = orc.generate(orc.gen_params(32 ** 3, 1, nlo=20000, nhi=20001,
                                      seed=2), with_vel=False)[0]
    grid = orc.cic(big, 64, xxmin=-0.5)
    rho = orc.cicm1(big, grid, xxmin=-0.5)
    _, isort, nnew = orc.tri_cic(rho)
    p = orc.greedy_params(32, ncut=1, mmin=20.0)
    r = orc.sod_greedy(big, np.arange(len(big)), isort, nnew, p)
    assert r["status"] == orc.E_UNCLOSED


def brute_centre(pos, nx, seed):
    """argmax over the particles of the 27 cells around the seed cell of
    sum_{d <= 1} 1/sqrt(d^2 + (1/64)^2), ties to the first in (cell walk,
    index) order -- written with numpy masks instead of cell lists."""
    xs = pos.astype(np.float64) * nx
    cell = np.floor(pos * np.float32(nx)).astype(int) % nx
    off = (cell - cell[seed] + nx // 2) % nx - nx // 2
    inner = np.all(np.abs(off) <= 1, axis=1)
    nb = np.all(np.abs(off) <= 2, axis=1)
    cand = np.where(inner)[0]
    rank = (off[cand, 0] + 1) * 9 + (off[cand, 1] + 1) * 3 + (off[cand, 2] + 1)
    order = np.lexsort((cand, rank))
    best, bs = -1, -1.0
    for c in cand[order]:
        d = np.abs(xs[nb] - xs[c])
        d = np.minimum(d, nx - d)
        d2 = (d * d).sum(axis=1)
        s = np.sum(1.0 / np.sqrt(d2[d2 <= 1.0] + (1.0 / 64) ** 2))
        if s > bs:
            best, bs = c, s
    return best, bs


def test_centre_search_brute_force(orc):
    nx = 16
    pos, vel, lab, ids, tab = orc.generate(
        orc.gen_params(nx ** 3, 12, nlo=50, nhi=600, seed=11))
    grid = orc.cic(pos, nx)
    rho = orc.cicm1(pos, grid)
    seeds, nmem = orc.fof_seeds(lab, rho, len(tab["size"]))
    cs, perm = orc.cell_sort(pos, nx, 8)
    r = orc.sod_fof(pos, cs, perm, 8, seeds, nmem,
                    orc.fof_params(nx, len(pos)))
    for h in range(len(seeds)):
        c, s = brute_centre(pos, nx, seeds[h])
        assert r["centre"][h] == c
        assert abs(r["s_best"][h] - s) <= 1e-12 * s


def test_fof_seeds_brute_force(orc):
    rng = np.random.default_rng(8)
    lab = rng.integers(-1, 5, size=400).astype(np.int32)
    rho = rng.integers(0, 7, size=400).astype(np.float64)  # ties
    seeds, nmem = orc.fof_seeds(lab, rho, 6)
    for h in range(6):
        m = np.where(lab == h)[0]
        assert nmem[h] == len(m)
        if len(m) == 0:
            assert seeds[h] == -1
        else:
            assert seeds[h] == m[np.argmax(rho[m])]  # first maximum


def test_isolated_nfw_r200(orc):
    nx, n = 32, 20000
    N = nx ** 3
    g = orc.gen_params(N, 1, nlo=n, nhi=n + 1, clo=6.0, chi=6.0, seed=4)
    pos, vel, lab, ids, tab = orc.generate(g)
    grid = orc.cic(pos, nx)
    rho = orc.cicm1(pos, grid)
    seeds, nmem = orc.fof_seeds(lab, rho, 1)
    cs, perm = orc.cell_sort(pos, nx, 8)
    r = orc.sod_fof(pos, cs, perm, 8, seeds, nmem, orc.fof_params(nx, N))
    mu = lambda x: np.log1p(x) - x / (1 + x)
    m200 = tab["size"][0] * mu(6.0) / mu(12.0)
    r200 = (m200 / (200.0 * 4.0 / 3.0 * np.pi)) ** (1.0 / 3.0)  # cells
    assert abs(r["rint"][0, 0] / r200 - 1.0) < 0.03
    assert abs(r["count"][0, 0] / m200 - 1.0) < 0.06


def test_profile_monotone_and_counts(orc):
    nx = 32
    pos, vel, lab, ids, tab = orc.generate(
        orc.gen_params(nx ** 3, 30, nlo=40, nhi=3000, seed=12))
    grid = orc.cic(pos, nx)
    rho = orc.cicm1(pos, grid)
    seeds, nmem = orc.fof_seeds(lab, rho, len(tab["size"]))
    cs, perm = orc.cell_sort(pos, nx, 8)
    r = orc.sod_fof(pos, cs, perm, 8, seeds, nmem,
                    orc.fof_params(nx, len(pos), rmax_frac=2.0), vel=vel)
    xs = pos.astype(np.float64) * nx
    for h in range(len(seeds)):
        prof = r["profile"][h]
        assert np.all(np.diff(prof) >= 0) and prof[0] >= 1
        c = r["centre"][h]
        d = np.abs(xs - xs[c])
        d = np.minimum(d, nx - d)
        rr = np.sqrt((d * d).sum(axis=1))
        R = r["radius"][h, 0]
        rj = R * np.exp(np.log(0.01) + np.arange(32) * (np.log(2.0) -
                                                         np.log(0.01)) / 31.0)
        rj[-1] = 2.0 * R  # the last radius is R * rmax_frac exactly
        np.testing.assert_array_equal(prof, (rr[None, :] < rj[:, None]).sum(1))
        inside = np.argsort(rr, kind="stable")[:r["count"][h, 0]]
        np.testing.assert_allclose(r["com_vel"][h], vel[inside].astype(
            np.float64).mean(axis=0), rtol=1e-12, atol=1e-9)


def test_computecom_closed_form(orc):
    # four particles straddling the periodic edge in x
    pos = np.array([[0.99, 0.5, 0.5], [0.01, 0.5, 0.5], [0.98, 0.6, 0.5],
                    [0.02, 0.4, 0.5]], np.float32)
    vel = np.array([[1, 0, 0], [3, 0, 0], [5, 2, 0], [7, -2, 0]], np.float32)
    cp, cv = orc.computecom(pos, vel)
    assert abs(min(cp[0], 1 - cp[0])) < 1e-6
    np.testing.assert_allclose(cp[1:], [0.5, 0.5], atol=1e-7)
    np.testing.assert_allclose(cv, [4.0, 0.0, 0.0])


def test_delta_conversions(orc):
    # Bryan & Norman (1998) at Omega_m = 0.27, z = 0: Delta_c = 97.0
    assert abs(orc.bryan_norman_mean(0.27) * 0.27 - 97.01) < 0.02
    assert abs(orc.bryan_norman_mean(0.27) - 359.3) < 0.1
    assert abs(orc.delta_crit_to_mean(200.0, 0.27) - 740.74) < 0.01
    assert abs(orc.bryan_norman_mean(1.0) - 18 * np.pi ** 2) < 1e-12


def halo_radii(z):
    """Sorted FP64 distances of the halo particles from the centre, from
    the stored float32 positions (the fixture's own numbers)."""
    nx = int(z["nx"])
    xs = z["pos"].astype(np.float64) * nx
    d = np.abs(xs[z["labels"] == 0] - xs[int(z["centre"])])
    d = np.minimum(d, nx - d)
    return np.sort(np.sqrt((d * d).sum(axis=1)))


def test_loglog_crossing_closed_form(orc):
    """rint/mint: O(k) is an exact power law A r^-2 across the crossing, so
    the log-log rule (haloanalyzer_new/src/modfunctions.f90:556-558) gives
    r* = (A/Delta)^(1/2) and M* = Delta (4 pi/3) r*^3 in closed form.  Only
    the float32 rounding of the positions (~3e-7 in r) separates them."""
    z = load("powerlaw_crossing.npz")
    r = fof_run(orc, z)
    assert r["centre"][0] == 0 and r["robust_centre"][0] == 1
    k = int(z["count"])
    assert r["count"][0, 0] == k
    assert abs(r["rint"][0, 0] / z["rint"] - 1) < 2e-6
    assert abs(r["mint"][0, 0] / z["mint"] - 1) < 6e-6
    # the same tolerances reject the plausible mistakes, computed from the
    # fixture's own radii around the crossing
    rs = halo_radii(z)
    r1, r2 = rs[k - 1], rs[k]
    assert r["radius"][0, 0] == r1
    V0, D = 4.0 / 3.0 * np.pi, float(z["delta"])
    o1, o2 = k / (V0 * r1 ** 3), (k + 1) / (V0 * r2 ** 3)
    linear = r1 + (D - o1) * (r2 - r1) / (o2 - o1)
    assert abs(linear / z["rint"] - 1) > 1e-4
    # the count of the outer particle off by one
    o2b = k / (V0 * r2 ** 3)
    off = np.exp(np.log(r1) + (np.log(D) - np.log(o1)) *
                 (np.log(r2) - np.log(r1)) / (np.log(o2b) - np.log(o1)))
    assert abs(off / z["rint"] - 1) > 1e-3
    # the discrete mass in place of the interpolated one
    assert abs(k / z["mint"] - 1) > 1e-2


def test_fof_com_across_periodic_edge(orc):
    """com_pos: symmetric pairs around a centre 1/16 cell from the x = 0
    face, every coordinate exact in float32: the centre of mass of the
    sphere is the centre, bit for bit.  Averaging without the minimum-image
    unwrap would put x near the middle of the box."""
    z = load("com_edge.npz")
    r = fof_run(orc, z)
    assert r["centre"][0] == 0 and r["robust_centre"][0] == 1
    assert r["count"][0, 0] == z["count"]
    np.testing.assert_array_equal(r["com_pos"][0], z["com_pos"])
    nx = int(z["nx"])
    xs = z["pos"].astype(np.float64) * nx
    d = np.abs(xs - xs[0])
    d = np.minimum(d, nx - d)
    inside = np.argsort(np.sqrt((d * d).sum(1)), kind="stable")
    inside = inside[:int(z["count"])]
    assert xs[inside, 0].max() - xs[inside, 0].min() > nx / 2  # straddles
    naive = xs[inside, 0].mean() / nx
    assert abs(naive - z["com_pos"][0]) > 0.1


def test_greedy_exclusion_of_overlapping_halos(orc):
    """Greedy exclusion (main file:2148-2166): halo A is grown first and
    removes its sphere; halo B, 0.45 cells away, grows from the alive
    particles only, so one of A's particles that lies inside B's radius is
    not counted: B's count is its design, 27 (28 if A's sphere were not
    excluded)."""
    z = load("greedy_pair.npz")
    pos, nx = z["pos"], int(z["nx"])
    grid = orc.cic(pos, 2 * nx, xxmin=-0.5)
    rho = orc.cicm1(pos, grid, xxmin=-0.5)
    _, isort, nnew = orc.tri_cic(rho)
    ids = np.arange(len(pos), dtype=np.int64)
    r = orc.sod_greedy(pos, ids, isort, nnew,
                       orc.greedy_params(nx, mmin=10.0, seuil2=800.0))
    assert r["status"] == orc.OK
    h = r["halos"]
    ka = int(np.where(h["centre"] == z["centre_a"])[0][0])
    kb = int(np.where(h["centre"] == z["centre_b"])[0][0])
    assert ka < kb  # A's seed is denser and is processed first
    assert h["masse"][ka] == z["count_a"] and h["masse"][kb] == z["count_b"]
    assert h["rayon"][ka] == z["radius_a"] and h["rayon"][kb] == z["radius_b"]
    assert h["accepted"][ka] == 1 and h["accepted"][kb] == 1
    # both spheres are excluded afterwards; B's sphere is exactly its own
    # particles
    assert r["alive"][z["sphere_a"]].sum() == 0
    assert r["alive"][z["sphere_b"]].sum() == 0
    np.testing.assert_array_equal(np.sort(np.where(r["structure"] ==
                                                   ids[z["centre_b"]])[0]),
                                  z["sphere_b"])
    # without the exclusion, B's sphere would hold one more particle
    xs = pos.astype(np.float64) * nx
    rb = np.sqrt(((xs - xs[int(z["centre_b"])]) ** 2).sum(1))
    extra = np.setdiff1d(np.where(rb <= float(z["radius_b"]))[0],
                         z["sphere_b"])
    assert len(extra) == 1 and extra[0] in z["sphere_a"]


def test_growth_at_given_centres_equals_search(orc):
    """or_sod_fof_at with the centres or_sod_fof found gives its outputs."""
    nx = 32
    pos, vel, lab, ids, tab = orc.generate(
        orc.gen_params(nx ** 3, 30, nlo=40, nhi=3000, seed=12))
    grid = orc.cic(pos, nx)
    rho = orc.cicm1(pos, grid)
    seeds, nmem = orc.fof_seeds(lab, rho, len(tab["size"]))
    cs, perm = orc.cell_sort(pos, nx, 8)
    p = orc.fof_params(nx, len(pos), delta2=800.0)
    a = orc.sod_fof(pos, cs, perm, 8, seeds, nmem, p, vel=vel)
    b = orc.sod_fof(pos, cs, perm, 8, seeds, nmem, p, vel=vel,
                    centres=a["centre"])
    for k in ("centre", "count", "radius", "rint", "mint", "count2",
              "profile", "com_pos", "com_vel", "nboundary", "w", "accepted"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
