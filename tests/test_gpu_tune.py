"""Tunables and the autotuner: every choice of the tunables gives the same
results (they only move halos and tiles between code paths that compute
the same values), and a tuning file written by the autotuner is read back
by a new context."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pass(ctx, nx=64, nclump=400, seed=21):
    import psod_amd as P
    N = nx ** 3
    g = P.gen_params(N, nclump, nlo=20, nhi=20000, seed=seed)
    pos, vel, lab, ids, tab = ctx.generate(g)
    nh = len(tab["size"])
    p = P.sod_params(nx, N, delta2=800.0, velocity=True)
    return P, p, pos, vel, lab, ids, nh


def test_tunables_do_not_change_results(ctx):
    P, p, pos, vel, lab, ids, nh = _pass(ctx)
    nx = p.nx
    base = ctx.tunables(nx)
    assert base == dict(small_max=1024, bound_min=1024, dense_tile=16384,
                        small_wg_per_cu=8, giant_min=1 << 20)
    runs = []
    for t in (dict(), dict(small_max=256, bound_min=64, dense_tile=512,
                           small_wg_per_cu=2, giant_min=2000),
              dict(small_max=768, bound_min=1 << 20, small_wg_per_cu=32,
                   giant_min=1)):
        ctx.set_tunables(**t)
        grid = ctx._empty((nx, nx, nx), ctx.torch.float32)
        o = ctx.run_pass(p, pos, lab, nh, vel=vel, ids=ids, grid=grid)
        runs.append(({k: v.cpu().numpy() for k, v in o.items()},
                     grid.cpu().numpy(), ctx.sod_stats()))
    ctx.set_tunables()
    assert ctx.tunables(nx) == base
    # the small-path cap moved halos between the one-wave and workgroup paths
    assert runs[1][2]["workgroup"] > runs[0][2]["workgroup"]
    o0, g0, _ = runs[0]
    assert np.all(o0["status"] & 1)
    for o, g, _ in runs[1:]:
        np.testing.assert_array_equal(g, g0)
        for k in ("centre", "centre_id", "count", "radius", "count2",
                  "profile", "com_pos", "nboundary"):
            np.testing.assert_array_equal(o[k], o0[k], err_msg=k)
        # the same flags except the path taken (PSOD_H_LARGE)
        np.testing.assert_array_equal(o["status"] & ~P.H_LARGE,
                                      o0["status"] & ~P.H_LARGE)
        for k in ("rint", "mint", "overdens", "overdens2"):
            np.testing.assert_allclose(o[k], o0[k], rtol=1e-13, err_msg=k)
        # the velocity sums follow the order of a path's walk
        np.testing.assert_allclose(o["com_vel"], o0["com_vel"], rtol=1e-12,
                                   atol=1e-12)


def test_growth_cubes_sorted_in_lds_give_the_binned_results(ctx, monkeypatch):
    """Workgroup halos whose growth cube holds at most PSOD_LIGHT_CAP
    particles (8192 unless lowered) are grown from their cube's keys sorted
    in LDS: the same outputs as the binned steps (cap 0), also with the two
    mixed (cap 1500).  The reach that closes a halo is the path's own."""
    P, p, pos, vel, lab, ids, nh = _pass(ctx, seed=5)
    runs = []
    try:
        # one-wave halos capped at 256 particles: more workgroup halos
        ctx.set_tunables(small_max=256)
        for cap in ("0", "1500", None):
            if cap is None:
                monkeypatch.delenv("PSOD_LIGHT_CAP", raising=False)
            else:
                monkeypatch.setenv("PSOD_LIGHT_CAP", cap)
            o = ctx.run_pass(p, pos, lab, nh, vel=vel, ids=ids)
            runs.append(({k: v.cpu().numpy() for k, v in o.items()},
                         ctx.sod_stats()))
    finally:
        monkeypatch.delenv("PSOD_LIGHT_CAP", raising=False)
        ctx.set_tunables()
    o0, s0 = runs[0]
    assert s0["light"] == 0 and s0["workgroup"] > 0
    assert 0 < runs[1][1]["light"] < runs[2][1]["light"]
    assert np.all(o0["status"] & 1)
    for o, _ in runs[1:]:
        for k in ("centre", "centre_id", "status", "count", "radius",
                  "count2", "profile", "com_pos", "nboundary"):
            np.testing.assert_array_equal(o[k], o0[k], err_msg=k)
        for k in ("rint", "mint", "overdens", "overdens2"):
            np.testing.assert_array_equal(o[k], o0[k], err_msg=k)
        np.testing.assert_allclose(o["com_vel"], o0["com_vel"], rtol=1e-12,
                                   atol=1e-12)


def test_tunables_reject_out_of_range(ctx):
    from psod_amd import PsodError
    with pytest.raises(PsodError):
        ctx.set_tunables(small_max=4096)
    with pytest.raises(PsodError):
        ctx.set_tunables(small_wg_per_cu=64)


def test_autotune_file_round_trip(ctx, tmp_path):
    import psod_amd as P
    _, p, pos, vel, lab, ids, nh = _pass(ctx, seed=3)
    f = str(tmp_path / "tune.txt")
    chosen, dep_ms, sod_ms = ctx.autotune(p, pos, lab, nh, vel=vel, ids=ids,
                                          reps=1, cache_file=f)
    assert dep_ms > 0 and sod_ms > 0
    assert ctx.tunables(p.nx) == chosen
    line = open(f).read().split()
    assert line[0] == "psod-tune" and int(line[3]) == p.nx
    assert [int(x) for x in line[6:11]] == [
        chosen["small_max"], chosen["bound_min"], chosen["dense_tile"],
        chosen["small_wg_per_cu"], chosen["giant_min"]]
    # a tuning file for another grid size leaves this one at its defaults
    with open(f, "a") as fh:
        fh.write(" ".join(line[:3] + ["128"] + line[4:6] +
                          ["512", "4096", "8192", "4", "65536", "1.0"]) + "\n")
    old = os.environ.get("PSOD_TUNE_FILE")
    os.environ["PSOD_TUNE_FILE"] = f
    try:
        c2 = P.Context(0)
    finally:
        if old is None:
            del os.environ["PSOD_TUNE_FILE"]
        else:
            os.environ["PSOD_TUNE_FILE"] = old
    try:
        assert c2.tunables(p.nx) == chosen
        assert c2.tunables(128) == dict(small_max=512, bound_min=4096,
                                        dense_tile=8192, small_wg_per_cu=4,
                                        giant_min=65536)
        assert c2.tunables(32)["small_max"] == 1024
    finally:
        c2.close()


def test_staged_growth_of_giant_halos(ctx, monkeypatch):
    """A 128^3 catalogue with a 1.5e6-particle central clump: halos grown
    by the staged launches (giant_min at or below their cubes; at 4096 more
    halos qualify than the 256 staged per pass, so the rest stay on one
    workgroup each) give the results of one workgroup per halo (giant_min
    above every cube), also with every band range of 2 or more keys sorted
    by the device-wide sort instead of the giant's workgroup."""
    import psod_amd as P
    nx = 128
    N = nx ** 3
    g = P.gen_params(N, 1500, nlo=20, nhi=20000, seed=17,
                     ncentral=1_500_000, ccentral=8.0)
    pos, vel, lab, ids, tab = ctx.generate(g)
    nh = len(tab["size"])
    p = P.sod_params(nx, N, deltas=(200.0, 500.0), delta2=800.0,
                     velocity=True)
    runs = []
    # every workgroup halo through the bins (the reach that closes a halo is
    # compared below, and cubes sorted whole in LDS may close at another)
    monkeypatch.setenv("PSOD_LIGHT_CAP", "0")
    try:
        for gm, smin in ((1 << 30, None), (1 << 20, None), (4096, None),
                         (4096, "2")):
            if smin:
                monkeypatch.setenv("PSOD_GIANT_SORT_MIN", smin)
            ctx.set_tunables(giant_min=gm)
            o = ctx.run_pass(p, pos, lab, nh, vel=vel, ids=ids)
            runs.append({k: v.cpu().numpy() for k, v in o.items()})
            st = ctx.sod_stats()
            assert (st["giants"] > 0) == (gm < 1 << 30), (gm, st)
    finally:
        monkeypatch.delenv("PSOD_GIANT_SORT_MIN", raising=False)
        monkeypatch.delenv("PSOD_LIGHT_CAP", raising=False)
        ctx.set_tunables()
    o0 = runs[0]
    assert np.all(o0["status"] & 1)
    assert o0["count"][:, 0].max() > 500_000
    for o in runs[1:]:
        for k in ("centre", "centre_id", "status", "w", "count", "radius",
                  "count2", "profile", "com_pos", "ngather", "nboundary"):
            np.testing.assert_array_equal(o[k], o0[k], err_msg=k)
        for k in ("rint", "mint", "overdens", "overdens2"):
            np.testing.assert_array_equal(o[k], o0[k], err_msg=k)
        np.testing.assert_allclose(o["com_vel"], o0["com_vel"], rtol=1e-12,
                                   atol=1e-12)
