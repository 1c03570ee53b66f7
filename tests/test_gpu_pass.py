"""GPU parity of the HIP pass (through the C ABI) against the FP64 oracle.

Bars: integer work (cell sort, counts, profiles, centre choice) bit-exact
except where the oracle itself flags a near-tie (a decision within 1e-5
relative of flipping under float32 rounding); float results within the
tolerances written in each test.
"""
import os

import numpy as np
import pytest

from parity import check_sod

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def catalogue(ctx, nx, nclump, nhi, seed, with_vel=True):
    import psod_amd as P
    N = nx ** 3
    g = P.gen_params(N, nclump, nlo=20, nhi=nhi, seed=seed)
    pos, vel, lab, ids, tab = ctx.generate(g, with_vel=with_vel)
    return dict(nx=nx, N=N, pos=pos, vel=vel, lab=lab, ids=ids, tab=tab,
                nh=len(tab["size"]))


@pytest.fixture(scope="module")
def c64(ctx):
    return catalogue(ctx, 64, 400, 20000, 9)


def host(t):
    return t.cpu().numpy()


def test_generator_matches_cpu_twin(ctx, orc):
    import psod_amd as P
    g = P.gen_params(32 ** 3, 80, nlo=20, nhi=5000, seed=3)
    pos, vel, lab, ids, tab = ctx.generate(g)
    opos, ovel, olab, oids, otab = orc.generate(
        orc.gen_params(32 ** 3, 80, nlo=20, nhi=5000, seed=3))
    np.testing.assert_array_equal(host(lab), olab)
    np.testing.assert_array_equal(host(ids), oids)
    hp = host(pos)
    # same draws, same FP64 formulas; device and host libm may differ in the
    # last ulp, which almost never survives the rounding to float32
    assert (hp == opos).mean() > 0.999
    np.testing.assert_allclose(hp, opos, rtol=0, atol=1e-6)
    assert hp.min() >= 0.0 and hp.max() < 1.0


def test_cell_sort_exact(ctx, orc, c64):
    nx = c64["nx"]
    cat = ctx.cell_sort(c64["pos"], nx, labels=c64["lab"])
    hp = host(c64["pos"])
    cs_o, perm_o = orc.cell_sort(hp, nx, 8)
    cs = host(cat["cell_start"]).astype(np.int64)
    np.testing.assert_array_equal(cs, cs_o)
    perm = host(cat["perm"])
    for c in range(nx ** 3):
        if cs[c + 1] - cs[c] > 1:
            perm[cs[c]:cs[c + 1]].sort()
    np.testing.assert_array_equal(perm, perm_o)
    rec = host(cat["rec"])
    idx = rec[:, 3].view(np.int32)
    np.testing.assert_array_equal(np.sort(idx), np.arange(len(hp)))
    # block-relative coordinates are exact: u + 8*block == x*nx
    xs = hp[idx] * np.float32(nx)
    blk = np.floor(xs / 8) * 8
    np.testing.assert_array_equal(rec[:, :3] + blk.astype(np.float32), xs)
    np.testing.assert_array_equal(host(cat["lab_sorted"]), host(c64["lab"])[idx])


def _morton9(rec):
    """Morton key of the 1/8-cell sub-cell of each record (the order the
    cell sort gives the records of a dense cell)."""
    q = (np.floor((rec[:, :3] - np.floor(rec[:, :3])) * 8).astype(np.int64)
         & 7)

    def spread(v):
        return (v & 1) | ((v & 2) << 2) | ((v & 4) << 4)
    return spread(q[:, 0]) | (spread(q[:, 1]) << 1) | (spread(q[:, 2]) << 2)


def test_dense_cells_in_morton_order(ctx):
    """Cells of at least 64 records, those sorted in registers (<= 4096)
    and those staged through a slice (> 4096), hold their records in Morton
    order of their 1/8-cell sub-cells: the centre search culls neighbour
    chunks by the compact boxes this gives."""
    import psod_amd as P
    nx = 128
    g = P.gen_params(nx ** 3, 200, nlo=20, nhi=3000, seed=3,
                     ncentral=1_500_000, ccentral=10.0)
    pos, _, lab, _, _ = ctx.generate(g, with_vel=False)
    cat = ctx.cell_sort(pos, nx, labels=lab)
    cs = host(cat["cell_start"]).astype(np.int64)
    key = _morton9(host(cat["rec"]))
    n = np.diff(cs)
    dense = np.where(n >= 64)[0]
    assert (n[dense] <= 4096).any() and (n[dense] > 4096).any(), n.max()
    for c in dense:
        k = key[cs[c]:cs[c + 1]]
        assert np.all(np.diff(k) >= 0), (c, n[c])


def test_cell_sort_wraps_unit_coordinate(ctx):
    import torch
    pos = torch.tensor([[1.0, 0.5, 0.25], [0.0, 0.999999, 0.5]],
                       device="cuda:0")
    cat = ctx.cell_sort(pos, 8)
    cs = host(cat["cell_start"])
    assert cs[-1] == 2
    # x == 1.0 lands in cell 0 like the reference's periodic wrap:
    # cells (0, 4, 2) and (0, 7, 4) of the single 8^3 block
    assert cs[161] - cs[160] == 1 and cs[313] - cs[312] == 1
    rec = host(cat["rec"])
    assert np.all(rec[:, :3] >= 0) and np.all(rec[:, :3] < 8)


@pytest.mark.parametrize("ratio,shift_frac", [(1, 0.0), (2, 0.0), (2, 0.5)])
def test_cic_deposit_and_interp(ctx, orc, c64, ratio, shift_frac):
    nx = c64["nx"]
    ng = ratio * nx
    shift = int(shift_frac * ng)
    cat = ctx.cell_sort(c64["pos"], nx)
    grid = ctx.cic_deposit(cat, ng, shift=shift)
    hp = host(c64["pos"])
    og = orc.cic(hp, ng, xxmin=-shift / ng)
    g = host(grid).astype(np.float64)
    # float32 weights (<= 1.8e-7 relative each), 2^-36 fixed-point sums,
    # float32 output: |err| <= 1e-6 |rho| + 1e-9
    assert np.all(np.abs(g - og) <= 1e-6 * np.abs(og) + 1e-9)
    m = float(np.float32(1.0 / len(hp))) * len(hp)
    assert abs(g.sum() / ng ** 3 - m) < 1e-6
    rho, _ = ctx.cic_interp(cat, grid, shift=shift)
    orho = orc.cicm1(hp, og, xxmin=-shift / ng)
    np.testing.assert_allclose(host(rho), orho, rtol=1e-6)


def test_seeds_ordered_after_torch_work(ctx, c64):
    """The seed keys are zeroed on torch's stream and filled by the
    library's stream.  Several milliseconds of matrix products queued on
    torch's stream delay the zeroing, so unordered streams would lose keys;
    the seeds must match those of an unloaded call."""
    import torch
    nx = c64["nx"]
    cat = ctx.cell_sort(c64["pos"], nx, labels=c64["lab"])
    grid = ctx.cic_deposit(cat, nx)
    ref = host(ctx.seeds(cat, grid, c64["nh"]))
    a = torch.rand((4096, 4096), device="cuda:0")
    for _ in range(8):
        a = torch.mm(a, a) * (1.0 / 4096)
    got = host(ctx.seeds(cat, grid, c64["nh"]))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(got, ref)
    lab = host(c64["lab"])
    assert np.all(lab[got] == np.arange(c64["nh"]))


def test_seeds_and_density_order(ctx, orc, c64):
    nx = c64["nx"]
    cat = ctx.cell_sort(c64["pos"], nx, labels=c64["lab"])
    grid = ctx.cic_deposit(cat, nx)
    seeds = host(ctx.seeds(cat, grid, c64["nh"]))
    hp = host(c64["pos"])
    orho = orc.cicm1(hp, orc.cic(hp, nx))
    oseeds, nmem = orc.fof_seeds(host(c64["lab"]), orho, c64["nh"])
    lab = host(c64["lab"])
    for h in np.where(seeds != oseeds)[0]:
        # only a near-tie in rho may pick another member
        assert lab[seeds[h]] == h
        assert orho[seeds[h]] >= orho[oseeds[h]] * (1 - 1e-6)
    assert (seeds == oseeds).mean() > 0.99
    rho, _ = ctx.cic_interp(cat, grid)
    isort, nnew = ctx.density_order(rho)
    r = host(rho)
    np.testing.assert_array_equal(host(isort) - 1,
                                  np.lexsort((np.arange(len(r)), -r)))
    assert nnew == int((r > 100.0).sum()) + 1


def sod_compare(ctx, orc, cat_d, deltas=(200.0,), delta2=800.0, nbins=32,
                rmin=0.01, rmax=2.0, plain=False):
    import psod_amd as P
    nx, N = cat_d["nx"], cat_d["N"]
    cat = ctx.cell_sort(cat_d["pos"], nx, labels=cat_d["lab"])
    grid = ctx.cic_deposit(cat, nx)
    seeds = ctx.seeds(cat, grid, cat_d["nh"])
    nmem = ctx.label_counts(cat_d["lab"], cat_d["nh"])
    p = P.sod_params(nx, N, deltas=deltas, delta2=delta2, nbins=nbins,
                     rmin_frac=rmin, rmax_frac=rmax, mmin=100.0,
                     velocity=True)
    o = ctx.sod_fof(p, cat, cat_d["pos"], seeds, nmem, vel=cat_d["vel"],
                    ids=cat_d["ids"], plain=plain)
    o = {k: host(v) for k, v in o.items()}
    hp = host(cat_d["pos"])
    cs, perm = orc.cell_sort(hp, nx, 8)
    fp = orc.fof_params(nx, N, deltas=deltas, delta2=delta2, nbins=nbins,
                        rmin_frac=rmin, rmax_frac=rmax, mmin=100.0)
    r = orc.sod_fof(hp, cs, perm, 8, host(seeds).astype(np.int64),
                    host(nmem).astype(np.int64), fp, vel=host(cat_d["vel"]),
                    ids=host(cat_d["ids"]))
    return o, r


def test_sod_fof_multi_delta(ctx, orc, c64):
    """The four thresholds, 64 bins on [0.005, 3] R of the multi-threshold
    workload."""
    deltas = (200.0, 500.0, 740.74, 359.3)
    o, r = sod_compare(ctx, orc, c64, deltas=deltas, nbins=64, rmin=0.005,
                       rmax=3.0)
    check_sod(o, r, 4)


def test_plain_fp64_matches_oracle(ctx, orc, c64):
    """The plain method is FP64 throughout: only the summation order of S
    differs from the oracle, so every result matches to the last bit
    except exact ties."""
    o, r = sod_compare(ctx, orc, c64, deltas=(200.0, 500.0), plain=True)
    assert np.all(o["status"] & 1)
    assert (o["centre"] == r["centre"]).mean() > 0.999
    same = o["centre"] == r["centre"]
    np.testing.assert_array_equal(o["count"][same, :2], r["count"][same, :2])
    np.testing.assert_array_equal(o["radius"][same, :2], r["radius"][same, :2])
    np.testing.assert_array_equal(o["profile"][same], r["profile"][same])
    np.testing.assert_array_equal(o["count2"][same], r["count2"][same])
    np.testing.assert_allclose(o["rint"][same, :2], r["rint"][same, :2],
                               rtol=1e-13)


def test_sod_retry_with_larger_slices(ctx, c64, monkeypatch):
    """Halos whose radial bands overflow a workgroup's collection slice are
    re-run with larger slices; the results do not depend on the slice."""
    import psod_amd as P
    nx, N = c64["nx"], c64["N"]
    cat = ctx.cell_sort(c64["pos"], nx, labels=c64["lab"])
    grid = ctx.cic_deposit(cat, nx)
    seeds = ctx.seeds(cat, grid, c64["nh"])
    nmem = ctx.label_counts(c64["lab"], c64["nh"])
    p = P.sod_params(nx, N, deltas=(200.0, 500.0), delta2=800.0,
                     velocity=True)

    def run():
        o = ctx.sod_fof(p, cat, c64["pos"], seeds, nmem, vel=c64["vel"],
                        ids=c64["ids"])
        return {k: host(v) for k, v in o.items()}, ctx.sod_stats()

    # every workgroup halo through the radial bins, whose bands fill the
    # slices (a cube sorted whole in LDS uses none)
    monkeypatch.setenv("PSOD_LIGHT_CAP", "0")
    a, sa = run()
    monkeypatch.setenv("PSOD_SLICE_KEYS", "16")
    b, sb = run()
    monkeypatch.delenv("PSOD_SLICE_KEYS")
    print("stats", sa, sb, flush=True)
    assert sa["retried"] == 0 and sb["retried"] > 0
    assert sb["workgroup"] == sa["workgroup"] > 0
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_staged_giants_retry_with_larger_slices(ctx, c64, monkeypatch):
    """Giant halos grown by the staged launches whose bands overflow their
    collection slice go to the re-run launch with larger slices; the
    results are those of one workgroup per halo with the default slice."""
    import psod_amd as P
    nx, N = c64["nx"], c64["N"]
    cat = ctx.cell_sort(c64["pos"], nx, labels=c64["lab"])
    grid = ctx.cic_deposit(cat, nx)
    seeds = ctx.seeds(cat, grid, c64["nh"])
    nmem = ctx.label_counts(c64["lab"], c64["nh"])
    p = P.sod_params(nx, N, deltas=(200.0, 500.0), delta2=800.0,
                     velocity=True)

    def run():
        o = ctx.sod_fof(p, cat, c64["pos"], seeds, nmem, vel=c64["vel"],
                        ids=c64["ids"])
        return {k: host(v) for k, v in o.items()}, ctx.sod_stats()

    # the one-workgroup results through the radial bins, as the staged
    # launches grow them (the reach that closes a halo is compared)
    monkeypatch.setenv("PSOD_LIGHT_CAP", "0")
    a, sa = run()
    try:
        ctx.set_tunables(giant_min=1)
        monkeypatch.setenv("PSOD_SLICE_KEYS", "16")
        b, sb = run()
    finally:
        monkeypatch.delenv("PSOD_SLICE_KEYS", raising=False)
        ctx.set_tunables()
    print("stats", sa, sb, flush=True)
    assert sa["giants"] == 0 and sb["giants"] > 0
    assert sa["retried"] == 0 and sb["retried"] > 0
    for k in a:
        if k == "com_vel":
            # the velocity sums follow the order of a path's walk
            np.testing.assert_allclose(a[k], b[k], rtol=1e-12, atol=1e-12)
        elif k in ("rint", "mint", "overdens", "overdens2"):
            np.testing.assert_allclose(a[k], b[k], rtol=1e-13, err_msg=k)
        else:
            np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_split_pilot_tasks_give_the_same_centres(ctx, c64, monkeypatch):
    """Pilot tasks split into pieces summed by separate waves (here every
    task with more than 64 records in reach, up to the split list's
    capacity) give the same integer sums, hence the same centres and
    results."""
    import psod_amd as P
    nx, N = c64["nx"], c64["N"]
    cat = ctx.cell_sort(c64["pos"], nx, labels=c64["lab"])
    grid = ctx.cic_deposit(cat, nx)
    seeds = ctx.seeds(cat, grid, c64["nh"])
    nmem = ctx.label_counts(c64["lab"], c64["nh"])
    p = P.sod_params(nx, N, deltas=(200.0, 500.0), delta2=800.0,
                     velocity=True)

    def run():
        o = ctx.sod_fof(p, cat, c64["pos"], seeds, nmem, vel=c64["vel"],
                        ids=c64["ids"])
        return {k: host(v) for k, v in o.items()}, ctx.sod_stats()

    a, sa = run()
    try:
        monkeypatch.setenv("PSOD_PILOT_BIG", "64")
        b, sb = run()
    finally:
        monkeypatch.delenv("PSOD_PILOT_BIG", raising=False)
    print("stats", sa, sb, flush=True)
    assert sa["centre_tasks_all"] == sb["centre_tasks_all"] > 0
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_whole_pass_matches_the_chain(ctx, c64):
    """psod_pass gives exactly the outputs of the step-by-step chain."""
    import psod_amd as P
    import torch
    nx, N, nh = c64["nx"], c64["N"], c64["nh"]
    p = P.sod_params(nx, N, deltas=(200.0, 500.0), delta2=800.0,
                     velocity=True)
    cat = ctx.cell_sort(c64["pos"], nx, labels=c64["lab"])
    grid = ctx.cic_deposit(cat, nx)
    seeds = ctx.seeds(cat, grid, nh)
    nmem = ctx.label_counts(c64["lab"], nh)
    a = ctx.sod_fof(p, cat, c64["pos"], seeds, nmem, vel=c64["vel"],
                    ids=c64["ids"])
    a = {k: host(v) for k, v in a.items()}
    g2 = torch.empty_like(grid)
    b = ctx.run_pass(p, c64["pos"], c64["lab"], nh, vel=c64["vel"],
                     ids=c64["ids"], grid=g2)
    b = {k: host(v) for k, v in b.items()}
    np.testing.assert_array_equal(host(grid), host(g2))
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_sod_fof_c1(ctx, orc):
    c1 = catalogue(ctx, 128, 2000, 20000, 1)
    o, r = sod_compare(ctx, orc, c1)
    check_sod(o, r, 1)
    # both execution paths were exercised
    assert (o["status"] & 4).any() and (~o["status"] & 4).any()
    # and the centre search's branch and bound: tasks bounded, boxes skipped
    # by their bound, boxes summed as items or abandoned part-way
    st = ctx.sod_stats()
    assert st["centre_tasks_all"] > 0 and st["pair_tests"] > 0
    assert st["centre_skipped"] > 0
    assert st["centre_aborted"] > 0


@pytest.mark.parametrize("name", ["shells_crossing.npz", "shells_floor.npz"])
@pytest.mark.parametrize("shift", [0.0, 8.0])
def test_sod_golden_shells(ctx, name, shift):
    import psod_amd as P
    import torch
    z = np.load(os.path.join(GOLD, name))
    nx = int(z["nx"])
    pos = z["pos"].astype(np.float64)
    # shift by whole cells so the halo straddles the periodic corner
    pos = ((pos * nx + shift) % nx / nx).astype(np.float32)
    pos[pos >= 1.0] = 0.0
    dpos = torch.from_numpy(pos).to("cuda:0")
    cat = ctx.cell_sort(dpos, nx)
    p = P.sod_params(nx, len(pos), deltas=(float(z["delta"]),))
    seeds = torch.tensor([int(z["centre"])], dtype=torch.int32,
                         device="cuda:0")
    nmem = torch.tensor([int((z["labels"] == 0).sum())], dtype=torch.int32,
                        device="cuda:0")
    o = {k: host(v) for k, v in ctx.sod_fof(p, cat, dpos, seeds,
                                             nmem).items()}
    assert o["centre"][0] == z["centre"]
    assert o["count"][0, 0] == z["count"]
    if shift == 0.0:
        assert abs(o["radius"][0, 0] - z["radius"]) <= 1e-12 * z["radius"]
        np.testing.assert_array_equal(o["profile"][0], z["profile"])
    else:
        # re-rounded positions move radii by float32 ulps only
        assert abs(o["radius"][0, 0] - z["radius"]) <= 1e-6 * z["radius"]


def test_empty_label_and_errors(ctx, c64):
    import psod_amd as P
    import torch
    nx = c64["nx"]
    cat = ctx.cell_sort(c64["pos"], nx)
    p = P.sod_params(nx, c64["N"])
    seeds = torch.tensor([-1, 5], dtype=torch.int32, device="cuda:0")
    o = {k: host(v) for k, v in ctx.sod_fof(p, cat, c64["pos"],
                                             seeds).items()}
    assert o["status"][0] == 0 and o["count"][0, 0] == 0
    assert o["status"][1] & 1
    with pytest.raises(P.PsodError) as e:
        ctx.cic_deposit(cat, 3 * nx)
    assert e.value.code == P.E_ARG
    bad = P.sod_params(nx, c64["N"], deltas=(500.0, 200.0))
    with pytest.raises(P.PsodError):
        ctx.sod_fof(bad, cat, c64["pos"], seeds)


def same_order_up_to_ties(isort, oisort, key):
    """isort equals oisort except inside runs of exactly equal keys, where
    the reference's quick_sort permutes ties its own way."""
    a = key[isort - 1]
    np.testing.assert_array_equal(a, key[oisort - 1])
    diff = np.where(isort != oisort)[0]
    for i in diff:
        assert a[i] == a[i - 1] or (i + 1 < len(a) and a[i] == a[i + 1])


@pytest.mark.parametrize("nx,nclump,xxmin", [(32, 40, -0.5), (32, 40, -0.3),
                                             (128, 2000, -0.5)])
def test_host_dropins_fp64(ctx, orc, nx, nclump, xxmin):
    """cic, cicm1 and tri_cic of the drop-in in FP64 against the oracle:
    the deposit differs only by the order of its FP64 additions, the
    interpolation of a given grid is the oracle's to the bit, and the
    density order of the oracle's densities is the oracle's order."""
    import psod_amd as P
    opos = orc.generate(orc.gen_params(nx ** 3, nclump, nlo=20, nhi=20000,
                                       seed=5), with_vel=False)[0]
    g = P.cic(opos, 2 * nx, xxmin=xxmin)
    og = orc.cic(opos, 2 * nx, xxmin=xxmin)
    assert g.dtype == np.float64
    # each node sums at most a few thousand terms: |err| < 1e-12 relative
    np.testing.assert_allclose(g, og, rtol=1e-12, atol=0)
    assert abs(g.sum() / (2 * nx) ** 3 - og.sum() / (2 * nx) ** 3) < 1e-12
    orho = orc.cicm1(opos, og, xxmin=xxmin)
    np.testing.assert_array_equal(P.cicm1(opos, og, xxmin=xxmin), orho)
    rh = P.cicm1(opos, g, xxmin=xxmin)
    np.testing.assert_allclose(rh, orho, rtol=1e-12)
    pdens, oisort, onn = orc.tri_cic(orho)
    isort, nn = P.tri_cic(orho)
    assert nn == onn
    same_order_up_to_ties(isort, oisort, 1.0 / orho)
    # the device entry point on resident densities gives the same order
    import torch
    disort, dnn = ctx.density_order_f64(torch.from_numpy(orho).to("cuda:0"))
    np.testing.assert_array_equal(host(disort), isort)
    assert dnn == nn
    with pytest.raises(P.PsodError):
        P.cic(opos, 2 * nx, mass=np.arange(len(opos), dtype=np.float32))


@pytest.mark.parametrize("nx,nclump,nhi,ncut", [(32, 60, 3000, 4),
                                                 (64, 400, 20000, 4)])
def test_spherover_greedy_matches_oracle(ctx, orc, nx, nclump, nhi, ncut):
    """Reference-semantics greedy SOD (seeds by density, exclusion) on the
    device against the oracle, record by record."""
    import psod_amd as P
    pos, vel, lab, ids, tab = orc.generate(
        orc.gen_params(nx ** 3, nclump, nlo=20, nhi=nhi, seed=21),
        with_vel=False)
    grid = orc.cic(pos, 2 * nx, xxmin=-0.5)
    rho = orc.cicm1(pos, grid, xxmin=-0.5)
    _, isort, nnew = orc.tri_cic(rho)
    r = orc.sod_greedy(pos, ids, isort, nnew,
                       orc.greedy_params(nx, ncut=ncut, ref_order=0))
    g = P.spherover(pos, ids, isort, nnew, nx, ncut=ncut)
    assert g["status"] == r["status"] == P.OK
    hr, hg = r["halos"], g["halos"]
    assert len(hg["ipart"]) == len(hr["ipart"]) > 10
    for k in ("ipart", "seed", "centre", "iamas", "masse", "masse2"):
        np.testing.assert_array_equal(hg[k], hr[k], err_msg=k)
    for k in ("rayon", "amass", "xc", "yc", "zc", "overdens"):
        np.testing.assert_array_equal(hg[k], hr[k], err_msg=k)
    np.testing.assert_array_equal(hg["accepted"], hr["accepted"])
    np.testing.assert_array_equal(g["mass_bins"], r["mass_bins"])
    np.testing.assert_allclose(g["rad_bins"], r["rad_bins"], rtol=1e-15)
    np.testing.assert_array_equal(g["alive"], r["alive"])
    np.testing.assert_array_equal(g["structure"], r["structure"])


def test_spherover_abort_paths(ctx, orc):
    import psod_amd as P
    z = np.load(os.path.join(GOLD, "shells_crossing.npz"))
    pos, nx = z["pos"], int(z["nx"])
    rho = orc.cicm1(pos, orc.cic(pos, 2 * nx, xxmin=-0.5), xxmin=-0.5)
    _, isort, nnew = orc.tri_cic(rho)
    ids = np.arange(len(pos), dtype=np.int64)
    g = P.spherover(pos, ids, isort, nnew, nx, mmin=20.0, nsel=50)
    assert g["status"] == P.E_CAPACITY
    g = P.spherover(pos, ids, isort, nnew, nx, mmin=20.0, seuil2=800.0)
    assert g["status"] == P.OK
    k = int(np.where(g["halos"]["centre"] == z["centre"])[0][0])
    assert g["halos"]["masse"][k] == z["count"]
    assert g["halos"]["masse2"][k] == 9
