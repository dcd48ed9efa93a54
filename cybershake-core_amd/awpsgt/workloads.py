"""The synthetic CyberShake-shaped workloads, built on the device.

* c1 -- the reference's CPU case: 200x200x100, h = 200 m, dt = 0.01 s, five
  layers (Vs 600 .. 3800 m/s), X force, 2000 SGT points;
* c2 -- the single-GPU site shape: 1500x1000x300 (450M points), h = 100 m,
  dt = 0.005 s, Vs = 500 -> 4500 m/s depth gradient times a seeded von Karman
  perturbation (awpsgt.media.c2_media), X force, 1.5M SGT points every 10
  steps;
* c3 -- 8 runs (4 sites x X/Y force) on grids from 400x300x150 to
  900x700x220, media as c2 per site (seeds 10..13), 200k SGT points each;
* c5 -- 16 runs (8 sites x X/Y) on grids from 500x400x150 to 1200x900x280
  (seeds 20..27), 200k SGT points each, admitted under a device-memory
  budget.

Every run has a Gaussian source time function (sigma 0.3 s, t0 1.2 s,
amplitude 1e9), NVAR 5 media (Qs = 0.05 Vs, Qp = 2 Qs), SOCALQ on, FL/FH/FP
0.01/25/0.5 Hz, a 20-cell Cerjan sponge (ARBC 0.92) and NTISKP_SGT 10.
"""
from __future__ import annotations

from typing import Callable, Dict, List, Optional, Tuple

import numpy as np

from . import AWP_STORAGE_FP32, In3d, Run, validate_in3d
from .batchq import site_grids
from .inputs import gaussian_stf, layered_media, random_cordfile_points

LAYERS_C1 = [600.0, 1400.0, 2200.0, 3000.0, 3800.0]

CONFIGS: Dict[str, Dict] = {
    "c2": dict(nx=1500, ny=1000, nz=300, dh=100.0, dt=0.005, nsgt=1_500_000, src=(750, 500, 10), seed=2,
               workload="single-GPU site shape 1500x1000x300 (450M points), h=100 m, dt=0.005 s, "
                        "von Karman Vs 500->4500 m/s, Qs=50 Vs_km, X force, 1.5M SGT points every 10 steps, FP32"),
    "c1": dict(nx=200, ny=200, nz=100, dh=200.0, dt=0.01, nsgt=2000, src=(100, 100, 10),
               workload="reference CPU case 200x200x100, h=200 m, dt=0.01 s, layered medium, "
                        "X force, 2000 SGT points every 10 steps, FP32"),
    "c3": dict(nsites=4, lo=(400, 300, 150), hi=(900, 700, 220), seed0=10, dh=100.0, dt=0.005, nsgt=200_000,
               workload="batched: 8 runs (4 sites x X/Y force), grids 400x300x150 .. 900x700x220, h=100 m, "
                        "dt=0.005 s, von Karman media per site, 200k SGT points per run, FP32"),
    "c5": dict(nsites=8, lo=(500, 400, 150), hi=(1200, 900, 280), seed0=20, dh=100.0, dt=0.005, nsgt=200_000,
               workload="batched queue: 16 runs (8 sites x X/Y force), grids 500x400x150 .. 1200x900x280, "
                        "h=100 m, dt=0.005 s, admitted by device-memory budget, 200k SGT points per run, FP32"),
}


def build_params(cfg: Dict, nt: int, nx: Optional[int] = None, ny: Optional[int] = None,
                 nz: Optional[int] = None, igreen: int = 1) -> In3d:
    """IN3D parameters of one run of a workload, nt time steps."""
    p = In3d()
    p.igreen = igreen
    p.dh, p.dt = cfg["dh"], cfg["dt"]
    p.tmax = (nt - 1) * cfg["dt"]
    p.npc, p.nd, p.arbc, p.pht = 0, 20, 0.92, 0.1
    p.nsrc, p.nst = 1, nt
    p.nx, p.ny, p.nz = nx or cfg["nx"], ny or cfg["ny"], nz or cfg["nz"]
    p.npx = p.npy = p.npz = 1
    p.ifault, p.mediarestart, p.nvar, p.socalq, p.nve = 0, 1, 5, 1, 1
    p.fl, p.fh, p.fp = 0.01, 25.0, 0.5
    p.read_step, p.ntiskp_sgt = nt, 10
    p.write_step = max(1, nt // 10)
    p.nt = nt
    validate_in3d(p)
    return p


def source_of(p: In3d, src: Tuple[int, int, int], amp: float = 1e9) -> Tuple[np.ndarray, np.ndarray]:
    """(xyz [1][3], stf [1][nt][6]): a Gaussian force along igreen's axis."""
    stf = np.zeros((1, p.nt, 6), dtype=np.float32)
    stf[0, :, p.igreen - 1] = gaussian_stf(p.nt, p.dt, 0.3, 1.2, amp)
    return np.array([src], dtype=np.int32), stf


def sgt_points(p: In3d, nsgt: int, seed: int) -> np.ndarray:
    return random_cordfile_points(nsgt, seed, (1, p.nx), (1, p.ny), (1, p.nz)).astype(np.int32)


def layered_nvar5(p: In3d) -> np.ndarray:
    m = layered_media(p.nx, p.ny, p.nz, LAYERS_C1)
    vs = m[..., 1]
    return np.stack([m[..., 0], vs, m[..., 2], 0.1 * vs, 0.05 * vs], -1).astype(np.float32)


def make_run(p: In3d, storage: int, media_fn: Callable, src: Tuple[int, int, int], nsgt: int, pts_seed: int,
             amp: float = 1e9, init: bool = True) -> Tuple[Run, int]:
    """A run with media_fn()'s media (numpy file layout, or a torch tensor on
    the device), a Gaussian force at src and nsgt seeded SGT points;
    initialised unless init is False.  Returns (run, points kept)."""
    import torch
    run = Run(p, device=0, storage=storage)
    nkept = run.set_sgt_points(sgt_points(p, nsgt, pts_seed))
    media = media_fn()
    if isinstance(media, np.ndarray):
        run.set_media(media)
    else:
        torch.cuda.synchronize()
        run.set_media_device(media.data_ptr())
        torch.cuda.synchronize()
    del media
    torch.cuda.empty_cache()
    run.set_source(*source_of(p, src, amp))
    if init:
        run.init()
    return run, nkept


def c2_run(nt: int, storage: int = AWP_STORAGE_FP32, amp: float = 1e9, nsgt: Optional[int] = None) -> Tuple[Run, int]:
    """The 1500x1000x300 site run with nt steps."""
    from .media import c2_media
    cfg = CONFIGS["c2"]
    p = build_params(cfg, nt)
    return make_run(p, storage, lambda: c2_media(p.nx, p.ny, p.nz, p.dh, seed=cfg["seed"]), cfg["src"],
                    cfg["nsgt"] if nsgt is None else nsgt, 3, amp)


def batch_jobs(config: str, nt: int) -> List[Dict]:
    """The runs of a batched workload (c3, c5) in queue order: site, seed,
    igreen and IN3D parameters of each."""
    cfg = CONFIGS[config]
    jobs = []
    for s, (nx, ny, nz) in enumerate(site_grids(cfg["nsites"], cfg["lo"], cfg["hi"])):
        for ig in (1, 2):
            jobs.append(dict(site=s, seed=cfg["seed0"] + s, igreen=ig,
                             p=build_params(cfg, nt, nx, ny, nz, igreen=ig)))
    return jobs


def batch_run(config: str, job: Dict, k: int, storage: int) -> Tuple[Run, int]:
    """Run k of a batched workload (job from batch_jobs), its media on the
    device, source at the grid centre 10 cells deep, SGT points seed 100 + k."""
    from .media import c2_media
    p = job["p"]
    return make_run(p, storage, lambda: c2_media(p.nx, p.ny, p.nz, p.dh, seed=job["seed"]),
                    (p.nx // 2, p.ny // 2, 10), CONFIGS[config]["nsgt"], 100 + k)
