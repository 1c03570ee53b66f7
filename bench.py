#!/usr/bin/env python3
"""Benchmark of the MI355X SOD pass.

One step = one pass of the hot path over the resident synthetic catalogue:
cell counting sort -> CIC deposit -> CIC^-1 + FoF seeds -> member counts ->
batched SOD of every halo (centre search, sphere growth to R_Delta,
profiles, centre of mass), replayed as one HIP graph (psod_pass_graph).  The
catalogue is generated on the device before the timed region.

Prints one JSON line (see the repository's measurement notes for the
definitions of every field).  Default workload: 1024^3 particles, 500,000
NFW clumps with N in [20, 2e6] drawn from N^-1.9, CIC grid 1024^3.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

CONFIGS = {
    # name: nx, clumps, N range, central clump, thresholds, profile
    "c1": dict(nx=128, nclump=2000, nhi=2e4),
    "c2": dict(nx=512, nclump=60000, nhi=2e5),
    "c3": dict(nx=1024, nclump=500000, nhi=2e6),
    "c4": dict(nx=1024, nclump=200000, nhi=1e5, ncentral=50_000_000,
               ccentral=6.0),
    "c5": dict(nx=1024, nclump=500000, nhi=2e6,
               deltas=(200.0, 500.0, 740.74, 359.3), nbins=64,
               rmin=0.005, rmax=3.0, velocity=True),
}
HBM_PEAK = 8.0e12     # MI355X, bytes/s
FP32_PEAK = 157.3e12  # MI355X vector FP32, flop/s


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3", choices=sorted(CONFIGS))
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--eager", action="store_true",
                    help="time eager passes (psod_pass) instead of graph "
                         "replays")
    ap.add_argument("--tune", action="store_true",
                    help="run the autotuner on the catalogue first")
    ap.add_argument("--giant-min", type=int, default=0,
                    help="growth cube particles from which a halo is grown by "
                         "the staged launches (0: the context's tunable)")
    ap.add_argument("--tunable", action="append", default=[],
                    metavar="NAME=VALUE",
                    help="set one of the context's tunables (repeatable)")
    ap.add_argument("--tune-file", default=None,
                    help="tuning file the autotuner appends to")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-greedy", action="store_true")
    ap.add_argument("--plain-sample", type=int, default=20000,
                    help="halos timed on the plain FP64 path (0: all)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--copy-check", action="store_true",
                    help="after the timed steps, copy the CIC grid once "
                         "(device to device, known size) to scale HBM "
                         "traffic counters")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpu():
    model = platform.processor() or "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count()
    return model, n


def subbox_oracle(ctx, pos, vel, seeds, nmem, grid, o, nx, p_kw, budget_s,
                  lo=None, sub=128, margin=16):
    """The FP64 oracle on one host core over a sub-box of the catalogue: its
    CIC deposit and the SOD of the halos seeded `margin`+ cells inside
    (members <= 20000), until the budget.  Returns the timing and the
    errors of the batched outputs against it (same seeds)."""
    import torch
    import oracle as O
    sub = min(sub, nx)
    lo = lo or [min(nx // 2, nx - sub)] * 3
    cell = torch.floor(pos * nx)
    inside = ((cell >= torch.tensor(lo, device=pos.device)) &
              (cell < torch.tensor(lo, device=pos.device) + sub)).all(dim=1)
    idx = torch.nonzero(inside).squeeze(1)
    del cell, inside
    lo_t = torch.tensor(lo, dtype=torch.float32, device=pos.device)
    spos = ((pos[idx] * nx - lo_t) / sub).cpu().numpy().astype(np.float32)
    sidx = idx.cpu().numpy()
    svel = vel[idx].cpu().numpy() if vel is not None else None
    N = pos.shape[0]
    m = np.full(len(spos), np.float32(1.0 / N), np.float32)
    t0 = time.perf_counter()
    og = O.cic(spos, sub, mass=m) * (nx / sub) ** 3
    t_cic = time.perf_counter() - t0
    # nodes 1..sub-1 of the sub-box get mass only from its own cells
    gg = grid[lo[2] + 1:lo[2] + sub, lo[1] + 1:lo[1] + sub,
              lo[0] + 1:lo[0] + sub].cpu().numpy().astype(np.float64)
    og = og[1:, 1:, 1:]
    nz = og > 1e-3
    cic_err = float((np.abs(gg - og)[nz] / og[nz]).max())
    sd = seeds.cpu().numpy()
    nm = nmem.cpu().numpy()
    pos_of = {int(j): i for i, j in enumerate(sidx)}
    cand = []
    for h in np.random.default_rng(0).permutation(len(sd)):
        s = int(sd[h])
        if s < 0 or nm[h] > 20000 or s not in pos_of:
            continue
        c = np.floor(spos[pos_of[s]] * sub)
        if np.all(c >= margin) and np.all(c < sub - margin):
            cand.append((h, pos_of[s], int(nm[h])))
    t0 = time.perf_counter()
    cs, perm = O.cell_sort(spos, sub, 8)
    fp = O.fof_params(sub, len(spos), **p_kw)
    fp.mpart = float(nx) ** 3 / float(N)  # the full box's mass unit
    res = []
    done = 0
    for k in range(0, len(cand), 8):
        chunk = cand[k:k + 8]
        r = O.sod_fof(spos, cs, perm, 8, np.array([c[1] for c in chunk]),
                      np.array([c[2] for c in chunk]), fp, vel=svel)
        res.append(([c[0] for c in chunk], r))
        done += len(chunk)
        if time.perf_counter() - t0 > budget_s:
            break
    t_sod = time.perf_counter() - t0
    # errors of the batched outputs on the same halos
    hs = np.concatenate([np.array(h) for h, _ in res])
    rr = {k: np.concatenate([r[k] for _, r in res]) for k in
          ("centre", "count", "radius", "rint", "mint", "margin", "rgap")}
    cen = sidx[rr["centre"]]
    oc = o["centre"].cpu().numpy()[hs]
    same = oc == cen
    nd = len(p_kw["deltas"])
    cnt = o["count"].cpu().numpy()[hs, :nd]
    rad = o["radius"].cpu().numpy()[hs, :nd]
    rint = o["rint"].cpu().numpy()[hs, :nd]
    mint = o["mint"].cpu().numpy()[hs, :nd]
    cmatch = (cnt == rr["count"][:, :nd]).all(axis=1) & same
    robust = (rr["margin"][:, :nd] > 1e-5).all(axis=1)
    errs = dict(
        halos=int(len(hs)), same_centre=int(same.sum()),
        same_count=int(cmatch.sum()),
        count_mismatch_robust=int((same & robust & ~cmatch).sum()),
        max_rel_err_R_discrete=float(np.abs(rad[cmatch] /
                                            rr["radius"][cmatch, :nd] - 1).max()),
        max_rel_err_R_interp=float(np.abs(rint[same] /
                                          rr["rint"][same, :nd] - 1).max()),
        max_rel_err_M_interp=float(np.abs(mint[same] /
                                          rr["mint"][same, :nd] - 1).max()),
        max_rel_err_cic_node=cic_err, cic_nodes=int(nz.sum()),
        sample=f"{sub}^3-cell sub-box at {lo} of the catalogue "
               f"({len(spos)} particles), halos seeded {margin}+ cells inside")
    return dict(value=done / t_sod, unit="halos/s",
                cic_gparticles_per_s=len(spos) / t_cic / 1e9,
                n=done, nsub=len(spos)), errs


def greedy_c1(ctx, budget_s):
    """The reference-semantics chain (cic -> cicm1 -> tri_cic -> greedy
    spherover, FP64) at the 128^3 workload: the drop-in on the GPU over all
    seeds, the FP64 oracle on one host core over the first seeds until the
    budget (records identical to the drop-in's first ones)."""
    import oracle as O
    import psod_amd as P
    nx = 128
    N = nx ** 3
    pos, _, _, ids, _ = ctx.generate(P.gen_params(N, 2000, nlo=20, nhi=2e4,
                                                  seed=1), with_vel=False)
    hp = pos.cpu().numpy()
    hid = ids.cpu().numpy()
    t0 = time.perf_counter()
    g = ctx.host_cic(hp, 2 * nx, xxmin=-0.5)
    rho = ctx.host_cicm1(hp, g, xxmin=-0.5)
    isort, nnew = ctx.host_tri_cic(rho)
    t_prep = time.perf_counter() - t0
    t0 = time.perf_counter()
    r = ctx.host_spherover(hp, hid, isort, nnew, nx)
    t_gpu = time.perf_counter() - t0
    nrec = len(r["halos"]["ipart"])
    nacc = int(r["halos"]["accepted"].sum())
    t0 = time.perf_counter()
    og = O.cic(hp, 2 * nx, xxmin=-0.5)
    orho = O.cicm1(hp, og, xxmin=-0.5)
    _, oisort, onnew = O.tri_cic(orho)
    t_cpu_prep = time.perf_counter() - t0
    same_order = bool(np.array_equal(oisort, isort)) and onnew == nnew
    # seeds processed in order; a prefix of them gives the first records
    k = 256
    while True:
        t0 = time.perf_counter()
        orr = O.sod_greedy(hp, hid, oisort, min(k, onnew),
                           O.greedy_params(nx, ref_order=0))
        t_cpu = time.perf_counter() - t0
        if t_cpu > budget_s / 4 or k >= onnew:
            break
        k *= 4
    ncpu = len(orr["halos"]["ipart"])
    match = bool(np.array_equal(orr["halos"]["iamas"],
                                r["halos"]["iamas"][:ncpu]) and
                 np.array_equal(orr["halos"]["masse"],
                                r["halos"]["masse"][:ncpu]))
    return dict(
        workload="128^3 particles, 2000 clumps on [20, 2e4], CIC grid 256^3 "
                 "(the reference's 2 nx), ncut 4, Delta 200, Delta2 800",
        gpu=dict(seeds=int(nnew), records=nrec, accepted=nacc,
                 prep_s=round(t_prep, 3), spherover_s=round(t_gpu, 3),
                 records_per_s=round(nrec / t_gpu, 1)),
        cpu=dict(kind="port", cores=1,
                 sample=f"FP64 oracle, 1 thread: greedy spherover over the "
                        f"first {min(k, onnew)} of {onnew} seeds "
                        f"({ncpu} records)",
                 prep_s=round(t_cpu_prep, 3), spherover_s=round(t_cpu, 3),
                 records_per_s=round(ncpu / t_cpu, 1)),
        density_order_identical=same_order, records_identical=match)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not "
              "supported (the pass is single-GPU by design)", file=sys.stderr)
        sys.exit(2)
    import torch
    import psod_amd as P

    cfg = CONFIGS[args.config]
    nx = cfg["nx"]
    N = nx ** 3
    ng = nx
    deltas = cfg.get("deltas", (200.0,))
    nbins = cfg.get("nbins", 32)
    vel_on = cfg.get("velocity", False)
    ctx = P.Context(0)
    g = P.gen_params(N, cfg["nclump"], nlo=20, nhi=cfg["nhi"],
                     seed=args.seed, ncentral=cfg.get("ncentral", 0),
                     ccentral=cfg.get("ccentral", 6.0))
    t0 = time.perf_counter()
    pos, vel, lab, ids, tab = ctx.generate(g, with_vel=vel_on)
    if not vel_on:
        vel = None
    nh = len(tab["size"])
    log(f"generated {N} particles, {nh} clumps "
        f"({tab['nclumped'] / N:.3f} clumped) in {time.perf_counter() - t0:.1f} s")
    p_kw = dict(deltas=deltas, delta2=800.0, nbins=nbins,
                rmin_frac=cfg.get("rmin", 0.01),
                rmax_frac=cfg.get("rmax", 2.0))
    p = P.sod_params(nx, N, velocity=vel_on, **p_kw)
    grid = torch.empty((ng, ng, ng), dtype=torch.float32, device="cuda:0")
    out = ctx.alloc_halo_out(nh, nbins)
    tuned = None
    if args.giant_min > 0:
        ctx.set_tunables(giant_min=args.giant_min)
    if args.tunable:
        cur = ctx.tunables(nx)
        for kv in args.tunable:
            k, v = kv.split("=")
            cur[k] = int(v)
        ctx.set_tunables(**cur)
    if args.tune:
        t0 = time.perf_counter()
        chosen, dep_ms, sod_ms = ctx.autotune(p, pos, lab, nh, vel=vel,
                                              ids=ids, reps=2,
                                              cache_file=args.tune_file)
        tuned = dict(chosen, deposit_ms=round(dep_ms, 3),
                     sod_ms=round(sod_ms, 3),
                     seconds=round(time.perf_counter() - t0, 1))
        log(f"autotuned in {tuned['seconds']} s: {chosen}")

    # stage times: eager steps (the same chain as one pass) with HIP events
    # on the library's stream
    ext = torch.cuda.ExternalStream(ctx.stream)
    names = ["sort", "deposit", "interp_seeds", "counts", "sod"]
    ev = {n: [] for n in names}

    def eager_step(record):
        e = [torch.cuda.Event(enable_timing=True) for _ in range(6)]
        if record:
            e[0].record(ext)
        cat = ctx.cell_sort(pos, nx, labels=lab, want_perm=False)
        if record:
            e[1].record(ext)
        ctx.cic_deposit(cat, ng, out=grid)
        if record:
            e[2].record(ext)
        seeds = ctx.seeds(cat, grid, nh)
        if record:
            e[3].record(ext)
        nmem = ctx.label_counts(cat["lab_sorted"], nh)
        if record:
            e[4].record(ext)
        ctx.sod_fof(p, cat, pos, seeds, nmem, vel=vel, ids=ids, out=out)
        if record:
            e[5].record(ext)
            for i, n in enumerate(names):
                ev[n].append((e[i], e[i + 1]))
        return cat, seeds, nmem

    for _ in range(args.warmup):
        eager_step(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        cat, seeds, nmem = eager_step(True)
    torch.cuda.synchronize()
    eager_ms = (time.perf_counter() - t0) / args.steps * 1e3
    cs = ctx.sod_stats()  # execution paths of the last eager step
    # per-stage median over the timed steps: a step that waits on a host-side
    # stall (seen now and then in later processes on a box, inside a stage's
    # events) does not move it
    st = {n: float(np.median([a.elapsed_time(b) for a, b in ev[n]]))
          for n in names}
    log(f"eager step {eager_ms:.2f} ms  stages(ms) " +
        " ".join(f"{k}={v:.3f}" for k, v in st.items()))

    # the timed steps: graph replays of the whole pass (or eager passes)
    def step():
        if args.eager:
            ctx.run_pass(p, pos, lab, nh, vel=vel, ids=ids, grid=grid,
                         out=out)
            return False
        return ctx.run_pass_graph(p, pos, lab, nh, vel=vel, ids=ids,
                                  grid=grid, out=out)[1]

    t0 = time.perf_counter()
    step()  # captures the graph
    capture_s = time.perf_counter() - t0
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    nrep = sum(bool(step()) for _ in range(args.steps))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = wall / args.steps * 1e3
    status = out["status"].cpu().numpy()
    ngath = out["ngather"].cpu().numpy().astype(np.int64)
    log(f"step {ms_step:.2f} ms ({'eager' if args.eager else 'graph'}, "
        f"{nrep}/{args.steps} replays, capture {capture_s:.2f} s)  "
        f"closed {(status & 1).sum()}/{nh}  large {(status & 4).sum() // 4}")

    # algorithmic bytes per launch (compulsory traffic only)
    b_dep = 12.0 * N + 4.0 * ng ** 3
    b_sort = 16.0 * N + 4.0 * (N + 1)
    b_interp = 12.0 * N + 4.0 * ng ** 3
    b_counts = 4.0 * N + 4.0 * nh
    b_sod = 16.0 * ngath.sum() + nh * (8 * 4 * 5 + 4 * nbins + 64)
    stages = {}
    for n, b in (("sort", b_sort), ("deposit", b_dep),
                 ("interp_seeds", b_interp), ("counts", b_counts),
                 ("sod", b_sod)):
        stages[n] = dict(ms=round(st[n], 4),
                         gbps=round(b / (st[n] * 1e-3) / 1e9, 1),
                         frac=round(b / (st[n] * 1e-3) / HBM_PEAK, 4))
    achieved = b_dep / (st["deposit"] * 1e-3)
    traffic = None
    tf = os.path.join(ROOT, "profiles", f"traffic_cic_deposit_{args.config}.json")
    if os.path.exists(tf):
        traffic = json.load(open(tf)).get("bytes_per_launch")
    roofline = dict(kernel="k_cic_deposit", bound="hbm",
                    achieved=round(achieved / 1e9, 1), peak=HBM_PEAK / 1e9,
                    unit="GB/s", frac=round(achieved / HBM_PEAK, 4),
                    traffic=traffic, algorithmic_bytes=b_dep,
                    launches_timed=args.steps)
    # centre search: pair tests against the vector FP32 peak (one pair test
    # is about 12 FP32 operations: 3 subtractions, 3 products and adds of the
    # squared distance, the compare, the reciprocal square root, the
    # fixed-point scale and the add)
    cms = cs["centre_ms"]
    pts = cs["pair_tests"] / (cms * 1e-3) if cms > 0 else 0.0
    centre = dict(
        pair_tests=cs["pair_tests"], ms=round(cms, 3),
        pair_tests_per_s=round(pts, 1),
        roofline=dict(bound="valu", flop_per_pair=12,
                      achieved=round(12 * pts / 1e12, 2),
                      peak=FP32_PEAK / 1e12, unit="TFLOP/s",
                      frac=round(12 * pts / FP32_PEAK, 4)),
        tasks=cs["centre_tasks_all"], units_evaluated=cs["centre_tasks"],
        units_skipped_by_bound=cs["centre_skipped"],
        groups_abandoned=cs["centre_aborted"])

    if args.copy_check:
        dst = torch.empty_like(grid)
        dst.copy_(grid)
        torch.mul(grid, 2.0, out=dst)  # elementwise kernel, same bytes
        torch.cuda.synchronize()
        log(f"copy check: {grid.numel() * 4} bytes read and written once "
            "(copy_, then mul)")
        del dst

    plain = None
    if not args.no_plain:
        sel = np.arange(nh)
        if args.plain_sample and args.plain_sample < nh:
            sel = np.sort(np.random.default_rng(0).choice(
                nh, args.plain_sample, replace=False))
        s_sel = seeds[torch.from_numpy(sel).to(seeds.device)]
        m_sel = nmem[torch.from_numpy(sel).to(seeds.device)]
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ctx.sod_fof(p, cat, pos, s_sel, m_sel, vel=vel, ids=ids, plain=True)
        torch.cuda.synchronize()
        tp = time.perf_counter() - t1
        t1 = time.perf_counter()
        ctx.sod_fof(p, cat, pos, s_sel, m_sel, vel=vel, ids=ids)
        torch.cuda.synchronize()
        tb = time.perf_counter() - t1
        plain = dict(halos=int(len(sel)), plain_halos_per_s=len(sel) / tp,
                     batched_halos_per_s_same_sample=len(sel) / tb,
                     speedup=tp / tb)
        log(f"plain FP64 per-halo: {len(sel)} halos in {tp:.2f} s; batched "
            f"same sample {tb * 1e3:.1f} ms; speedup {tp / tb:.1f}x")

    model, ncores = host_cpu()
    cpu = errors = greedy = None
    if not args.no_cpu:
        # the graph's outputs, grid and seeds are those of the last replay
        sb, errors = subbox_oracle(ctx, pos, vel, seeds, nmem, grid, out, nx,
                                   p_kw, args.cpu_seconds)
        cpu = dict(value=sb["value"], unit="halos/s", cores=1, kind="port",
                   host_cpu=model, host_threads=ncores,
                   sample=(f"FP64 oracle, 1 thread: cell sort + SOD of "
                           f"{sb['n']} halos (<= 20000 members) of a "
                           f"128^3-cell sub-box ({sb['nsub']} particles) of "
                           f"the {nx}^3 catalogue; CIC of the sub-box at "
                           f"{sb['cic_gparticles_per_s']:.4f} Gparticles/s"),
                   cic_gparticles_per_s=sb["cic_gparticles_per_s"])
        log(f"cpu baseline {cpu['value']:.1f} halos/s; errors {errors}")
        if not args.no_greedy:
            greedy = greedy_c1(ctx, args.cpu_seconds)
            greedy["cpu"]["host_cpu"] = model
            log(f"greedy c1: {greedy}")

    line = {
        "metric": "SOD halos/s + CIC Gparticles/s at 1024^3, 1 MI355X, vs "
                  "plain FP64 GPU",
        "value": round(nh / (ms_step * 1e-3), 1),
        "unit": "halos/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "precision": "f32 positions split as int cell + f32 offset; f64 "
                     "radii, overdensities and COM; i64 fixed-point sums "
                     "(CIC weights, centre-search terms)",
        "data": "synthetic (seeded NFW clumps on a uniform background, "
                "generated on the device)",
        "config": {"workload": f"{args.config}: {N} particles ({nx}^3), "
                               f"{nh} FoF halos, CIC grid {ng}^3",
                   "particles": N, "halos": nh, "cic_grid": ng,
                   "deltas": list(deltas), "profile_bins": nbins,
                   "step": "eager pass" if args.eager else
                           "one HIP graph replay of the whole pass",
                   "parallelism": "single GPU"},
        "graph": dict(replays_timed=nrep, capture_s=round(capture_s, 3),
                      eager_ms_per_step=round(eager_ms, 3)),
        "tunables": tuned or ctx.tunables(nx),
        "cic_gparticles_per_s": round(N / (st["deposit"] * 1e-3) / 1e9, 2),
        "sod_stage_halos_per_s": round(nh / (st["sod"] * 1e-3), 1),
        "stages": stages,
        "roofline": roofline,
        "centre_search": centre,
        "sod_paths": dict(wave=cs["wave"], workgroup=cs["workgroup"],
                          workgroup_sorted_in_lds=cs["light"],
                          rerun=cs["retried"], giants_staged=cs["giants"],
                          giant_cycles=cs["giant_cycles"]),
        "plain_fp64": plain,
        "cpu_baseline": cpu,
        "errors_vs_fp64_oracle": errors,
        "greedy_c1": greedy,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
