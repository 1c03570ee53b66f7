#!/usr/bin/env python3
"""Benchmark of the MI355X SOD pass.

One step = one pass of the hot path over the resident synthetic catalogue:
cell counting sort -> CIC deposit -> CIC^-1 + FoF seeds -> batched SOD of
every halo (centre search, sphere growth to R_Delta, profiles, centre of
mass).  The catalogue is generated on the device before the timed region.

Prints one JSON line (see the repository's measurement notes for the
definitions of every field).  Default workload: 1024^3 particles, 500,000
NFW clumps with N in [20, 2e6] drawn from N^-1.9, CIC grid 1024^3.
"""
from __future__ import annotations

import argparse
import json
import os
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
HBM_PEAK = 8.0e12  # MI355X, bytes/s


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3", choices=sorted(CONFIGS))
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
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


def cpu_baseline(ctx, pos, lab, seeds, nmem, nx, budget_s):
    """The FP64 oracle on one host core over a 128^3-cell sub-box of the
    catalogue: its CIC deposit and the SOD of the halos whose seeds lie
    16+ cells inside the sub-box (members <= 20000), until the budget."""
    import torch
    import oracle as O
    sub, x0, margin = 128, nx // 2, 16
    cell = torch.floor(pos * nx)
    inside = ((cell >= x0) & (cell < x0 + sub)).all(dim=1)
    idx = torch.nonzero(inside).squeeze(1)
    spos = ((pos[idx] * nx - x0) / sub).cpu().numpy().astype(np.float32)
    sidx = idx.cpu().numpy()
    remap = {int(j): i for i, j in enumerate(sidx)}
    t0 = time.perf_counter()
    O.cic(spos, sub)
    t_cic = time.perf_counter() - t0
    sd = seeds.cpu().numpy()
    nm = nmem.cpu().numpy()
    cand = []
    for h in np.random.default_rng(0).permutation(len(sd)):
        s = int(sd[h])
        if s < 0 or nm[h] > 20000 or s not in remap:
            continue
        c = np.floor(spos[remap[s]] * sub)
        if np.all(c >= margin) and np.all(c < sub - margin):
            cand.append((remap[s], int(nm[h])))
    t0 = time.perf_counter()
    cs, perm = O.cell_sort(spos, sub, 8)
    p = O.fof_params(sub, len(spos))
    p.mpart = float(nx) ** 3 / float(len(pos))  # the full box's mass unit
    done = 0
    for k in range(0, len(cand), 8):
        chunk = cand[k:k + 8]
        O.sod_fof(spos, cs, perm, 8, np.array([c[0] for c in chunk]),
                  np.array([c[1] for c in chunk]), p)
        done += len(chunk)
        if time.perf_counter() - t0 > budget_s:
            break
    t_sod = time.perf_counter() - t0
    return dict(value=done / t_sod, unit="halos/s", cores=1, kind="port",
                sample=(f"FP64 oracle, 1 thread: cell sort + SOD of {done} "
                        f"halos (<= 20000 members, seeds 16+ cells inside) of "
                        f"a 128^3-cell sub-box ({len(spos)} particles) of "
                        f"the {nx}^3 catalogue; CIC of the sub-box at "
                        f"{len(spos) / t_cic / 1e9:.4f} Gparticles/s"),
                cic_gparticles_per_s=len(spos) / t_cic / 1e9)


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
    nh = len(tab["size"])
    log(f"generated {N} particles, {nh} clumps "
        f"({tab['nclumped'] / N:.3f} clumped) in {time.perf_counter() - t0:.1f} s")
    nmem = ctx.label_counts(lab, nh)
    p = P.sod_params(nx, N, deltas=deltas, delta2=800.0, nbins=nbins,
                     rmin_frac=cfg.get("rmin", 0.01),
                     rmax_frac=cfg.get("rmax", 2.0), velocity=vel_on)
    grid = torch.empty((ng, ng, ng), dtype=torch.float32, device="cuda:0")
    out = ctx.alloc_halo_out(nh, nbins)
    ext = torch.cuda.ExternalStream(ctx.stream)
    names = ["sort", "deposit", "interp_seeds", "sod"]
    ev = {n: [] for n in names}

    def step(record):
        e = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
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
        ctx.sod_fof(p, cat, pos, seeds, nmem, vel=vel, ids=ids, out=out)
        if record:
            e[4].record(ext)
            for i, n in enumerate(names):
                ev[n].append((e[i], e[i + 1]))
        return cat, seeds

    for _ in range(args.warmup):
        step(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        cat, seeds = step(True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = wall / args.steps * 1e3
    cs = ctx.sod_stats()  # execution paths of the last timed step
    st = {n: float(np.mean([a.elapsed_time(b) for a, b in ev[n]]))
          for n in names}
    status = out["status"].cpu().numpy()
    ngath = out["ngather"].cpu().numpy().astype(np.int64)
    log(f"step {ms_step:.2f} ms  stages(ms) " +
        " ".join(f"{k}={v:.3f}" for k, v in st.items()) +
        f"  closed {(status & 1).sum()}/{nh}  large {(status & 4).sum()}")

    # algorithmic bytes per launch (compulsory traffic only)
    b_dep = 12.0 * N + 4.0 * ng ** 3
    b_sort = 16.0 * N + 4.0 * (N + 1)
    b_interp = 12.0 * N + 4.0 * ng ** 3
    b_sod = 16.0 * ngath.sum() + nh * (8 * 4 * 5 + 4 * nbins + 64)
    stages = {}
    for n, b in (("sort", b_sort), ("deposit", b_dep),
                 ("interp_seeds", b_interp), ("sod", b_sod)):
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
                    traffic=traffic, algorithmic_bytes=b_dep)

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

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(ctx, pos, lab, seeds, nmem, nx, args.cpu_seconds)
        log(f"cpu baseline {cpu['value']:.1f} halos/s")

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
                   "parallelism": "single GPU"},
        "cic_gparticles_per_s": round(N / (st["deposit"] * 1e-3) / 1e9, 2),
        "sod_stage_halos_per_s": round(nh / (st["sod"] * 1e-3), 1),
        "stages": stages,
        "roofline": roofline,
        "centre_search": dict(
            pair_tests=cs["pair_tests"], ms=round(cs["centre_ms"], 3),
            pair_tests_per_s=(round(cs["pair_tests"] / (cs["centre_ms"] * 1e-3), 1)
                              if cs["centre_ms"] > 0 else None),
            tasks=cs["centre_tasks_all"], units_evaluated=cs["centre_tasks"],
            units_skipped_by_bound=cs["centre_skipped"],
            groups_abandoned=cs["centre_aborted"]),
        "sod_paths": dict(wave=cs["wave"], workgroup=cs["workgroup"],
                          rerun=cs["retried"]),
        "plain_fp64": plain,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
