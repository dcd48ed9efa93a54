#!/usr/bin/env python3
"""Throughput of the AWP-ODC-SGT time step on one MI355X.

Workload (default, "c2"): the single-GPU CyberShake site shape -- grid
1500x1000x300 (450M points), h = 100 m, dt = 0.005 s; Vs = 500 -> 4500 m/s
depth gradient times a seeded von Karman perturbation (correlation 5 km,
Hurst 0.1, 5% std, clipped to +-3 sigma), Vp = 1.73 Vs, Gardner density,
Qs = 50 Vs[km/s], Qp = 2 Qs; X point force with a Gaussian source time
function; 1.5M SGT points sampled every 10 steps.  Synthetic inputs are built
on the device; the timed region is K full time steps with every input
already resident in HBM.  A step is the two-stage one (velocity kernel,
free-surface images, stress kernel with sponge and memory variables, source,
SGT gather) or the one-sweep fused kernel in place of the first three: by
default (--scheme auto) the run takes the library's default scheme for its
storage mode and awp_tune keeps the fastest scheme and launch shape before
the warm-up; --scheme split / fused fixes the scheme at its default launch.

Batched workloads: "c3" -- 8 runs (4 sites x X/Y force) on grids from
400x300x150 to 900x700x220, 200k SGT points each; "c5" -- 16 runs (8 sites x
X/Y) on grids from 500x400x150 to 1200x900x280, admitted in waves under the
device-memory budget.  Same media recipe per site (seeded), h = 100 m,
dt = 0.005 s.  Each wave is one awpsgt.Batch (one persistent launch per stage
for all its runs, replayed from a HIP graph), tuned before the timed steps.

    python bench.py [--steps K] [--warmup W] [--config c2|c1|c3|c5] [--storage fp32|mixed]
                    [--scheme auto|split|fused] [--compare-fp32]

Prints one JSON line.  value = point-updates per second of the whole job
(Gpt-updates/s).  roofline = the dominant kernel: the stress kernel of the
two-stage step (124 B FP32 / 92 B mixed per point: every array it touches,
once, at its storage width), or the fused kernel (the whole step's 176 B /
142 B -- every array each of the reference's two stages touches -- although
the sweep itself must move only 140 B / 106 B); achieved = those bytes x
points / the kernel's average time from HIP events on the run's stream.  The
176 / 142 B per point-update are the canonical algorithmic bytes of the step
(what the two reference stages touch), so `frac` compares the fused and the
two-stage step on one basis; `frac_vs_fused_min_bytes` states the fused
kernel against its own 140 / 106 B.  Batched configs report the whole step
(all kernels of a wave) at 176 / 142 B.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "cybershake-core_amd"))

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per point-update of the two-stage path: velocity stage
# read+write u,v,w (24), read 6 stresses (24) and d1 (4 FP32 / 2 FP16); stress
# stage read u,v,w (12), read+write 6 stresses (48), read+write r1..r6 (48
# FP32 / 24 FP16), read lam, mu (8 / 4) and qp, qs (8 / 4)
BYTES_UPDATE = {"fp32": 176, "mixed": 142}
BYTES_STR = {"fp32": 124, "mixed": 92}     # the two-stage stress kernel alone
# what the fused sweep must move: read+write u,v,w, 6 stresses and r1..r6,
# read d1, lam, mu, qp, qs once (140 FP32; mixed 106)
BYTES_FUSED = {"fp32": 140, "mixed": 106}

from awpsgt.workloads import CONFIGS, build_params, make_run  # noqa: E402


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)

# SGT error of the storage modes against the FP64 restatement, per component
# (rel L2 over the whole series set; 800x600x200, 1500 steps, 2000 points:
# tests/test_gpu_c4.py measures and bounds them)
SGT_ERROR_VS_FP64 = {
    "fp32": {"rel_l2": [1.8e-6, 2.7e-6], "peak": [1.6e-6, 2.7e-6]},
    "mixed": {"rel_l2": [1.7e-4, 3.1e-4], "peak": [1.1e-4, 3.0e-4]},
    "case": "800x600x200, h=200 m, 1500 steps, layered medium, 2000 SGT points; range over the six components",
}


def cpu_leg(a):
    if a.no_cpu_baseline:
        return None
    sys.path.insert(0, REPO)
    from oracle import cpu_baseline
    # the GPU lease's share of the host is 16 cores, whatever the affinity
    # mask shows; both are stated
    aff = len(os.sched_getaffinity(0))
    ncores = min(aff, 16)
    log("cpu baseline on %d cores (%d in the affinity mask) ..." % (ncores, aff))
    out = cpu_baseline.measure(ncores)
    out["cores_in_affinity_mask"] = aff
    out["cores_note"] = "MPI ranks = the 16-core host share of one GPU's lease"
    return out


def traffic_of(name):
    tfile = os.path.join(REPO, "profiles", name)
    if os.path.exists(tfile):
        for ln in open(tfile):
            if ln.startswith("hbm_bytes_per_launch="):
                return float(ln.split("=", 1)[1])
    return None


def bench_single(a, storage):
    import numpy as np
    import torch
    import awpsgt
    from awpsgt.media import c2_media
    from awpsgt.workloads import layered_nvar5

    cfg = CONFIGS[a.config]
    nt = a.warmup + a.steps
    p = build_params(cfg, nt)
    npts_grid = p.nx * p.ny * p.nz
    log("%s: grid %dx%dx%d, %d warmup + %d timed steps" % (a.config, p.nx, p.ny, p.nz, a.warmup, a.steps))
    t0 = time.time()
    if a.config == "c2":
        media_fn = lambda: c2_media(p.nx, p.ny, p.nz, p.dh, seed=cfg["seed"])  # noqa: E731
    else:
        media_fn = lambda: layered_nvar5(p)  # noqa: E731
    st = awpsgt.AWP_STORAGE_MIXED if storage == "mixed" else awpsgt.AWP_STORAGE_FP32
    run, nkept = make_run(p, st, media_fn, cfg["src"], cfg["nsgt"], 3)
    log("set-up %.1f s, %d SGT points, %.1f GB on device" % (time.time() - t0, nkept, run.device_bytes() / 1e9))
    if a.scheme == "auto":
        t1 = time.time()
        run.tune(3)
        log("tuned in %.1f s" % (time.time() - t1))
    fused, tile_y, nchunk = run.scheme()
    log("time step: %s, 64 x %d tiles, %d z chunks" % ("fused one-sweep" if fused else "two-stage", tile_y, nchunk))

    run.step(a.warmup)
    run.synchronize()
    torch.cuda.synchronize()
    run.set_timing(True)
    t = time.perf_counter()
    run.step(a.steps)
    run.synchronize()
    torch.cuda.synchronize()
    el = time.perf_counter() - t
    tv, ts, ta = run.timing()
    step_now, nsamp = run.progress()
    assert step_now == nt
    finite = bool(np.isfinite(run.sgt()).all())
    run.close()

    value = npts_grid * a.steps / el / 1e9
    ms_step = el / a.steps * 1e3
    kern_ms = ts / a.steps                   # fused kernel, or the stress kernel of the two-stage step
    if fused:
        kname = ("k_step_batch (fused velocity + stress + sponge + free surface sweep over the run's one-entry "
                 "job table; persistent blocks claiming tiles per XCD; k_step under AWP_CLAIM=0)")
        bytes_pt = BYTES_UPDATE[storage]
        tname = "step_%s%s_traffic.txt" % (a.config, "" if storage == "fp32" else "_" + storage)
    else:
        kname = "k_stress_batch (stress + memory variables + sponge + free surface; the run's one-entry job table)"
        bytes_pt = BYTES_STR[storage]
        tname = "stress_%s%s_traffic.txt" % (a.config, "" if storage == "fp32" else "_" + storage)
    ach = bytes_pt * npts_grid / (kern_ms * 1e-3) / 1e9
    log("%.3f ms/step; kernel %.3f ms (%.0f GB/s algorithmic at %d B/pt); finite=%s"
        % (ms_step, kern_ms, ach, bytes_pt, finite))
    roof = {"bound": "hbm", "kernel": kname, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic_of(tname),
            "algorithmic_bytes_per_point": bytes_pt, "algorithmic_bytes_per_launch": bytes_pt * npts_grid,
            "avg_launch_ms": round(kern_ms, 4),
            "step_frac": round(value * BYTES_UPDATE[storage] / HBM_PEAK_GBS, 4)}
    if fused:
        # the same time against the bytes the fused sweep itself must move
        # (each array read once and written once): the kernel's own fraction
        roof["fused_kernel_min_bytes_per_point"] = BYTES_FUSED[storage]
        roof["frac_vs_fused_min_bytes"] = round(ach * BYTES_FUSED[storage] / bytes_pt / HBM_PEAK_GBS, 4)
    wl = cfg["workload"] if storage == "fp32" else cfg["workload"].replace(
        ", FP32", ", mixed storage (FP16 r1..r6 with per-row exponents, FP16 qp/qs, "
        "FP16 d1/mu/lam deviations from level profiles)")
    config = {"workload": wl, "grid": [p.nx, p.ny, p.nz], "points": npts_grid, "sgt_points": int(nkept),
              "storage": storage, "step": "fused" if fused else "two-stage",
              "launch": {"tile_rows": tile_y, "z_chunks": nchunk, "tuned": a.scheme == "auto"},
              "parallelism": "single GPU"}
    return value, ms_step, roof, config


def bench_batched(a, storage, config=None, steps=None, warmup=None):
    """One batched workload (c3 / c5) in one storage mode: runs admitted in
    waves under the budget, each wave one tuned awpsgt.Batch; returns
    (Gpt-updates/s, ms per step, roofline, config)."""
    import numpy as np
    import torch
    import awpsgt
    from awpsgt.batchq import Need, admit
    from awpsgt.workloads import batch_jobs, batch_run

    config = config or a.config
    steps = steps or a.steps
    warmup = a.warmup if warmup is None else warmup
    cfg = CONFIGS[config]
    nt = warmup + steps
    st = awpsgt.AWP_STORAGE_MIXED if storage == "mixed" else awpsgt.AWP_STORAGE_FP32
    jobs = batch_jobs(config, nt)
    for j in jobs:
        j["need"] = Need(*awpsgt.estimate_bytes(j["p"], st, cfg["nsgt"]))
    free, _ = torch.cuda.mem_get_info()
    budget = int(a.budget_gb * 1e9) if a.budget_gb else free - (8 << 30)
    waves = admit([j["need"] for j in jobs], budget)
    log("%s %s: %d runs, %.1f GB steady in total, budget %.1f GB -> %d wave(s) of %s runs"
        % (config, storage, len(jobs), sum(j["need"].steady for j in jobs) / 1e9, budget / 1e9, len(waves),
           "/".join(str(len(w)) for w in waves)))
    tot_pts = 0
    tot_time = 0.0
    kern = []
    launch = None
    fused = False
    persist = True
    for w in waves:
        runs = []
        t0 = time.time()
        for k in w:
            r, _ = batch_run(config, jobs[k], k, st)
            runs.append(r)
        pts = sum(r.p.nx * r.p.ny * r.p.nz for r in runs)
        log("wave of %d runs, %.0fM points, set-up %.1f s, %.1f GB on device"
            % (len(runs), pts / 1e6, time.time() - t0, sum(r.device_bytes() for r in runs) / 1e9))
        with awpsgt.Batch(runs) as b:
            if not a.no_tune:
                t1 = time.time()
                launch = b.tune(2)
                log("tuned in %.1f s: %s, %s grids, tile rows %d, %s %d, z chunks %d (%.2f ms per step's kernels)"
                    % (time.time() - t1, "fused" if b.fused() else "two-stage",
                       "persistent" if b.scheme()[1] else "one-block-per-tile", launch[0],
                       "prefetch" if b.fused() else "z unroll", launch[1], launch[2], b.time_ms()))
            else:
                launch = b.launch()
            fused, persist = b.scheme()
            b.step(warmup)
            torch.cuda.synchronize()
            t = time.perf_counter()
            b.step(steps)
            el = time.perf_counter() - t
            kern.append(b.time_ms() / steps)
        finite = all(bool(np.isfinite(r.sgt()).all()) for r in runs)
        for r in runs:
            assert r.progress()[0] == nt
            r.close()
        torch.cuda.empty_cache()
        log("  %.3f ms/step (device %.3f), %.2f Gpt/s, finite=%s" % (el / steps * 1e3, kern[-1],
                                                                    pts * steps / el / 1e9, finite))
        tot_pts += pts * steps
        tot_time += el
    value = tot_pts / tot_time / 1e9
    pts_step = tot_pts / steps
    ms_step = tot_time / steps * 1e3
    dev_ms = sum(kern)
    ach = BYTES_UPDATE[storage] * pts_step / (dev_ms * 1e-3) / 1e9
    # measured bytes per step of the two-stage batch (the fused batch has
    # none measured: traffic null)
    tname = "batch_%s_%s%s_traffic.txt" % (config, storage, "_fused" if fused else "")
    roof = {"bound": "hbm",
            "kernel": ("whole batched step: k_step_batch (fused sweep over every run of a wave) + sources/SGT"
                       if fused else
                       "whole batched step: k_velocity_batch + k_fs_velocity_batch + k_stress_batch (every run "
                       "of a wave) + sources/SGT"),
            "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
            "traffic": traffic_of(tname), "algorithmic_bytes_per_point": BYTES_UPDATE[storage],
            "algorithmic_bytes_per_step": int(BYTES_UPDATE[storage] * pts_step),
            "device_ms_per_step": round(dev_ms, 4)}
    if fused:
        roof["fused_kernel_min_bytes_per_point"] = BYTES_FUSED[storage]
        roof["frac_vs_fused_min_bytes"] = round(ach * BYTES_FUSED[storage] / BYTES_UPDATE[storage] / HBM_PEAK_GBS, 4)
    wl = cfg["workload"] if storage == "fp32" else cfg["workload"].replace(
        ", FP32", ", mixed storage (FP16 r1..r6 with per-row exponents, FP16 qp/qs, "
        "FP16 d1/mu/lam deviations from level profiles)")
    conf = {"workload": wl, "runs": len(jobs), "waves": [len(w) for w in waves],
            "points_per_step": int(pts_step), "storage": storage,
            "step": ("fused" if fused else "two-stage") + ", one launch per stage over all runs, HIP graph",
            "launch": {"tile_rows": launch[0], ("prefetch" if fused else "z_unroll"): launch[1],
                       "z_chunks": launch[2], "persistent_grid": persist} if launch else None,
            "budget_gb": round(budget / 1e9, 1), "parallelism": "single GPU"}
    return value, ms_step, roof, conf


def batched_record(a):
    """The metric's second half: the batched workload in FP32 and in mixed
    storage, each run created with both wavefield buffers where the budget
    allows so that the batch tuner picks the faster scheme for each storage
    mode (the comparison is best against best), their ratio and the SGT
    error each storage mode is stated at."""
    cfg = a.batched
    old = os.environ.get("AWP_SCHEME")
    os.environ["AWP_SCHEME"] = "fused"
    try:
        v32, ms32, roof32, c32 = bench_batched(a, "fp32", cfg, a.batched_steps, 2)
        vmx, msmx, roofmx, cmx = bench_batched(a, "mixed", cfg, a.batched_steps, 2)
    finally:
        if old is None:
            os.environ.pop("AWP_SCHEME", None)
        else:
            os.environ["AWP_SCHEME"] = old
    log("batched %s: mixed %.3f Gpt/s vs FP32 %.3f Gpt/s: x%.3f" % (cfg, vmx, v32, vmx / v32))
    return {"config": cfg, "workload": CONFIGS[cfg]["workload"].replace(", FP32", ""),
            "steps": a.batched_steps, "warmup": 2,
            "fp32": {"value": round(v32, 4), "unit": "Gpt-updates/s", "ms_per_step": round(ms32, 4),
                     "step": c32["step"], "launch": c32["launch"], "waves": c32["waves"], "roofline": roof32},
            "mixed": {"value": round(vmx, 4), "unit": "Gpt-updates/s", "ms_per_step": round(msmx, 4),
                      "step": cmx["step"], "launch": cmx["launch"], "waves": cmx["waves"], "roofline": roofmx},
            "mixed_vs_fp32": round(vmx / v32, 4),
            "target_mixed_vs_fp32": 1.6,
            "bytes_ceiling_mixed_vs_fp32": {"two_stage": round(BYTES_UPDATE["fp32"] / BYTES_UPDATE["mixed"], 3),
                                             "fused": round(BYTES_FUSED["fp32"] / BYTES_FUSED["mixed"], 3)},
            "sgt_error_vs_fp64": SGT_ERROR_VS_FP64}


def claim_ab(a):
    """One run of the single-grid workload stepped alternately with the
    claimed launches and with the static ones (AWP_CLAIM=2 / 0, switched
    between steps on the same buffers), so that both forms see the same
    memory.  Prints one JSON line; a diagnostic, not the throughput line."""
    import torch
    import awpsgt
    from awpsgt.media import c2_media
    from awpsgt.workloads import layered_nvar5

    cfg = CONFIGS[a.config]
    rounds = 3
    p = build_params(cfg, rounds * 2 * (a.warmup + a.steps))
    if a.config == "c2":
        media_fn = lambda: c2_media(p.nx, p.ny, p.nz, p.dh, seed=cfg["seed"])  # noqa: E731
    else:
        media_fn = lambda: layered_nvar5(p)  # noqa: E731
    st = awpsgt.AWP_STORAGE_MIXED if a.storage == "mixed" else awpsgt.AWP_STORAGE_FP32
    run, _ = make_run(p, st, media_fn, cfg["src"], cfg["nsgt"], 3)
    fused = run.scheme()[0]
    out = []
    for r in range(rounds):
        for c in ("2", "0"):
            os.environ["AWP_CLAIM"] = c
            run.set_scheme(fused)
            run.step(a.warmup)
            run.synchronize()
            t = time.perf_counter()
            run.step(a.steps)
            run.synchronize()
            ms = (time.perf_counter() - t) / a.steps * 1e3
            log("claim=%s: %.3f ms/step" % (c, ms))
            out.append({"claim": c != "0", "ms_per_step": round(ms, 4)})
    run.close()
    torch.cuda.synchronize()
    print(json.dumps({"claim_ab": out, "config": a.config, "storage": a.storage,
                      "step": "fused" if fused else "two-stage"}))


def class_probe(a):
    """Runs of the single-grid workload created and freed in a fixed order
    (A; B while A lives; C after A is freed; D after B is freed), each timed
    over a few steps: a run's speed follows the physical memory its
    allocation gets, and a run created into memory another run freed runs
    like that run.  Prints one JSON line; a diagnostic, not the throughput
    line."""
    import torch
    import awpsgt
    from awpsgt.media import c2_media
    from awpsgt.workloads import layered_nvar5

    cfg = CONFIGS[a.config]
    p = build_params(cfg, 4 * (a.warmup + a.steps))
    if a.config == "c2":
        media = c2_media(p.nx, p.ny, p.nz, p.dh, seed=cfg["seed"])
    else:
        media = layered_nvar5(p)
    st = awpsgt.AWP_STORAGE_MIXED if a.storage == "mixed" else awpsgt.AWP_STORAGE_FP32
    out = []

    def make():
        return make_run(p, st, lambda: media, cfg["src"], cfg["nsgt"], 3)[0]

    def timed(r, name):
        r.step(a.warmup)
        r.synchronize()
        t = time.perf_counter()
        r.step(a.steps)
        r.synchronize()
        ms = (time.perf_counter() - t) / a.steps * 1e3
        log("%s: %.3f ms/step" % (name, ms))
        out.append({"run": name, "ms_per_step": round(ms, 4)})

    ra = make()
    timed(ra, "A")
    rb = make()
    timed(rb, "B (A alive)")
    timed(ra, "A again")
    ra.close()
    torch.cuda.synchronize()
    rc = make()
    timed(rc, "C (after A freed)")
    rb.close()
    torch.cuda.synchronize()
    rd = make()
    timed(rd, "D (after B freed)")
    rc.close()
    rd.close()
    print(json.dumps({"class_probe": out, "config": a.config, "storage": a.storage}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--storage", default="fp32", choices=["fp32", "mixed"])
    ap.add_argument("--scheme", default="auto", choices=["auto", "split", "fused"],
                    help="time-step scheme of the runs: auto = single runs take the library default and "
                         "awp_tune, batches take two-stage runs; fused batches tune over both schemes")
    ap.add_argument("--compare-fp32", action="store_true",
                    help="with --storage mixed: also run the FP32 workload and report the ratio")
    ap.add_argument("--budget-gb", type=float, default=0.0,
                    help="device-memory budget for the batched configs (default: free memory - 8 GiB)")
    ap.add_argument("--no-tune", action="store_true", help="batched configs: skip the tuner")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batched", default="c5", choices=["c3", "c5", "none"],
                    help="with the single-grid line: also run this batched workload in FP32 and in mixed "
                         "storage and add their ratio as the 'batched' sub-record")
    ap.add_argument("--batched-steps", type=int, default=10)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device copy of 4 GiB five times (a known-size kernel "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--claim-ab", action="store_true",
                    help="only step one single-grid run alternately with claimed and static tile launches "
                         "(same buffers) and print their times")
    ap.add_argument("--class-probe", action="store_true",
                    help="only create, time and free single-grid runs in a fixed order (the speed a run gets "
                         "follows the memory its allocation gets) and print their times")
    a = ap.parse_args()
    if a.class_probe:
        if a.scheme != "auto":
            os.environ["AWP_SCHEME"] = a.scheme
        class_probe(a)
        return
    if a.claim_ab:
        if a.scheme != "auto":
            os.environ["AWP_SCHEME"] = a.scheme
        claim_ab(a)
        return
    if a.copy_calibration:
        import torch
        x = torch.ones(1 << 30, device="cuda", dtype=torch.float32)
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 << 30, "copy_bytes_written": 4 << 30, "launches": 5}))
        return
    if a.gpus != 1:
        print("bench.py measures one GPU: the solver runs one SGT job (or one batch of jobs) per GPU "
              "and this build has no multi-GPU path (--gpus %d requested)" % a.gpus, file=sys.stderr)
        sys.exit(2)

    batched = a.config in ("c3", "c5")
    if a.scheme == "auto":
        # batches: two-stage runs (the fused step's second wavefield buffer
        # would admit fewer runs per wave)
        if batched:
            os.environ["AWP_SCHEME"] = "split"
        else:
            os.environ.pop("AWP_SCHEME", None)
    else:
        os.environ["AWP_SCHEME"] = a.scheme
    bench_fn = bench_batched if batched else bench_single
    value, ms_step, roof, config = bench_fn(a, a.storage)
    if a.compare_fp32 and a.storage == "mixed":
        v32, ms32, _, _ = bench_fn(a, "fp32")
        config["fp32_value"] = round(v32, 4)
        config["mixed_vs_fp32"] = round(value / v32, 4)
        log("mixed %.3f Gpt/s vs FP32 %.3f Gpt/s: x%.3f" % (value, v32, value / v32))
    sub = None
    if not batched and a.batched != "none":
        sub = batched_record(a)
    cpu = cpu_leg(a)
    out = {
        "metric": "Gpt-updates/s + % HBM roofline; batched mixed-storage vs FP32 at stated SGT error",
        "value": round(value, 4),
        "unit": "Gpt-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": config,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if sub is not None:
        out["batched"] = sub
    print(json.dumps(out))


if __name__ == "__main__":
    main()
