#!/usr/bin/env python3
"""Throughput of the AWP-ODC-SGT time step on one MI355X.

Workload (default, "c2"): the single-GPU CyberShake site shape -- grid
1500x1000x300 (450M points), h = 100 m, dt = 0.005 s; Vs = 500 -> 4500 m/s
depth gradient times a seeded von Karman perturbation (correlation 5 km,
Hurst 0.1, 5% std, clipped to +-3 sigma), Vp = 1.73 Vs, Gardner density,
Qs = 50 Vs[km/s], Qp = 2 Qs; X point force with a Gaussian source time
function; 1.5M SGT points sampled every 10 steps.  Synthetic inputs are built
on the device; the timed region is K full time steps (velocity, free surface,
stress + memory variables, sponge, source, SGT gather) with every input
already resident in HBM.

    python bench.py [--steps K] [--warmup W] [--config c2|c1] [--storage fp32|mixed]

Prints one JSON line.  value = point-updates per second of the whole job
(Gpt-updates/s); roofline = the stress kernel (the dominant one): its
algorithmic bytes per launch (FP32: 124 B per point, 31 FP32 arrays touched
once; mixed: 96 B per point, r1..r6 BF16 and qp/qs FP16) over its average
launch time from HIP events on the run's stream.
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
BYTES_VEL = 52       # per point-update, velocity stage (13 FP32 arrays, both storage modes)
# per point-update, stress stage: read u,v,w (12 B); read+write 6 stresses
# (48 B); read+write r1..r6 (48 B FP32, 24 B BF16); read lam, mu (8 B) and
# qp, qs (8 B FP32, 4 B FP16)
BYTES_STR = {"fp32": 124, "mixed": 96}

CONFIGS = {
    "c2": dict(nx=1500, ny=1000, nz=300, dh=100.0, dt=0.005, nsgt=1_500_000, src=(750, 500, 10),
               workload="single-GPU site shape 1500x1000x300 (450M points), h=100 m, dt=0.005 s, "
                        "von Karman Vs 500->4500 m/s, Qs=50 Vs_km, X force, 1.5M SGT points every 10 steps, FP32"),
    "c1": dict(nx=200, ny=200, nz=100, dh=200.0, dt=0.01, nsgt=2000, src=(100, 100, 10),
               workload="reference CPU case 200x200x100, h=200 m, dt=0.01 s, layered medium, "
                        "X force, 2000 SGT points every 10 steps, FP32"),
}


def build_params(cfg, nt):
    import awpsgt
    p = awpsgt.In3d()
    p.igreen = 1
    p.dh, p.dt = cfg["dh"], cfg["dt"]
    p.tmax = (nt - 1) * cfg["dt"]
    p.npc, p.nd, p.arbc, p.pht = 0, 20, 0.92, 0.1
    p.nsrc, p.nst = 1, nt
    p.nx, p.ny, p.nz = cfg["nx"], cfg["ny"], cfg["nz"]
    p.npx = p.npy = p.npz = 1
    p.ifault, p.mediarestart, p.nvar, p.socalq, p.nve = 0, 1, 5, 1, 1
    p.fl, p.fh, p.fp = 0.01, 25.0, 0.5
    p.read_step, p.ntiskp_sgt = nt, 10
    p.write_step = max(1, nt // 10)
    p.nt = nt
    awpsgt.validate_in3d(p)
    return p


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--storage", default="fp32", choices=["fp32", "mixed"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device copy of 4 GiB five times (a known-size kernel "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE counters)")
    a = ap.parse_args()
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
        print("bench.py measures one GPU: the solver runs one SGT job per GPU and this build "
              "has no multi-GPU path (--gpus %d requested)" % a.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import awpsgt
    from awpsgt.inputs import gaussian_stf, random_cordfile_points, layered_media
    from awpsgt.media import c2_media

    cfg = CONFIGS[a.config]
    nt = a.warmup + a.steps
    p = build_params(cfg, nt)
    npts_grid = p.nx * p.ny * p.nz
    print("[bench] %s: grid %dx%dx%d, %d warmup + %d timed steps" % (a.config, p.nx, p.ny, p.nz, a.warmup, a.steps),
          file=sys.stderr, flush=True)

    t0 = time.time()
    storage = awpsgt.AWP_STORAGE_MIXED if a.storage == "mixed" else awpsgt.AWP_STORAGE_FP32
    run = awpsgt.Run(p, device=0, storage=storage)
    if a.config == "c2":
        media = c2_media(p.nx, p.ny, p.nz, p.dh, seed=2)
        torch.cuda.synchronize()
        run.set_media_device(media.data_ptr())
        del media
        torch.cuda.empty_cache()
    else:
        m = layered_media(p.nx, p.ny, p.nz, [600.0, 1400.0, 2200.0, 3000.0, 3800.0])
        vs = m[..., 1]
        m5 = np.stack([m[..., 0], vs, m[..., 2], 0.1 * vs, 0.05 * vs], -1).astype(np.float32)
        run.set_media(m5)
    stf = np.zeros((1, nt, 6), dtype=np.float32)
    stf[0, :, 0] = gaussian_stf(nt, p.dt, 0.3, 1.2, 1e9)
    run.set_source(np.array([cfg["src"]], dtype=np.int32), stf)
    pts = random_cordfile_points(cfg["nsgt"], 3, (1, p.nx), (1, p.ny), (1, p.nz))
    nkept = run.set_sgt_points(pts.astype(np.int32))
    run.init()
    print("[bench] set-up %.1f s, %d SGT points, %.1f GB on device"
          % (time.time() - t0, nkept, run.device_bytes() / 1e9), file=sys.stderr, flush=True)

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
    s = run.sgt()
    finite = bool(np.isfinite(s).all())
    run.close()

    value = npts_grid * a.steps / el / 1e9
    ms_step = el / a.steps * 1e3
    str_ms = ts / a.steps
    vel_ms = tv / a.steps
    bstr = BYTES_STR[a.storage]
    str_gbs = bstr * npts_grid / (str_ms * 1e-3) / 1e9
    vel_gbs = BYTES_VEL * npts_grid / (vel_ms * 1e-3) / 1e9
    print("[bench] %.3f ms/step; velocity %.3f ms (%.0f GB/s), stress %.3f ms (%.0f GB/s); finite=%s"
          % (ms_step, vel_ms, vel_gbs, str_ms, str_gbs, finite), file=sys.stderr, flush=True)

    traffic = None
    tfile = os.path.join(REPO, "profiles", "stress_%s%s_traffic.txt"
                         % (a.config, "" if a.storage == "fp32" else "_" + a.storage))
    if os.path.exists(tfile):
        for ln in open(tfile):
            if ln.startswith("hbm_bytes_per_launch="):
                traffic = float(ln.split("=", 1)[1])

    cpu = None
    if not a.no_cpu_baseline:
        sys.path.insert(0, REPO)
        from oracle import cpu_baseline
        ncores = min(len(os.sched_getaffinity(0)), 16)
        print("[bench] cpu baseline on %d cores ..." % ncores, file=sys.stderr, flush=True)
        cpu = cpu_baseline.measure(ncores)

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
        "config": {"workload": cfg["workload"] if a.storage == "fp32" else
                   cfg["workload"].replace(", FP32", ", mixed storage (BF16 r1..r6, FP16 qp/qs)"), "grid": [p.nx, p.ny, p.nz], "points": npts_grid,
                   "sgt_points": int(nkept), "storage": a.storage, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": "k_stress (stress + memory variables + sponge + free surface)",
                     "achieved": round(str_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(str_gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes_per_launch": bstr * npts_grid,
                     "avg_launch_ms": round(str_ms, 4),
                     "step_frac": round(value * (BYTES_VEL + bstr) / HBM_PEAK_GBS, 4)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
