"""CPU baseline timing for bench.py (test infrastructure: the cpu_baseline
leg is the only product-side caller, and only to time the checker).

Preferred: the reference CPU solver itself (oracle/_ref/pmcl3d, built from
the reference Fortran by oracle/Makefile.ref) run with MPICH's mpiexec on the
host cores.  The grid is 400x400x200 (32M points, h = 100 m, dt = 0.005 s,
the layered medium of the 200x200x100 reference case), large enough that each
rank's arrays (>= 2M points, >= 160 MB) do not sit in the host caches, as
they do not at the single-GPU site shape.  Two runs of different length are
timed by wall clock and differenced, so start-up and media reading cancel
(the reference's own per-step line prints only 3 decimals).
Fallback when that binary or mpiexec is absent: the OpenMP C restatement
(oracle/awp_oracle.c) on the same grid.
"""
from __future__ import annotations

import os
import re
import shutil
import subprocess
import sys
import tempfile
import time
from typing import Dict

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
REF = os.path.join(HERE, "_ref", "pmcl3d")
MPIEXEC = "/opt/conda/bin/mpiexec"
sys.path.insert(0, os.path.join(REPO, "cybershake-core_amd"))
from awpsgt.inputs import RunSpec, layered_media, random_cordfile_points, write_run  # noqa: E402

LAYERS = [600.0, 1400.0, 2200.0, 3000.0, 3800.0]


NX, NY, NZ = 400, 400, 200


def _decomp(cores: int):
    """Largest NPX x NPY x NPZ <= cores with >= 2 ranks per horizontal axis,
    each rank at least 20 cells wide."""
    best = (2, 2, 1)
    for px in (2, 4, 5):
        for py in (2, 4, 5):
            for pz in (1, 2, 4, 5):
                if px * py * pz <= cores and px * py * pz > best[0] * best[1] * best[2]:
                    best = (px, py, pz)
    return best


def _spec(nsteps: int, dec, nx=NX, ny=NY, nz=NZ, dh=100.0, dt=0.005) -> RunSpec:
    return RunSpec(nx=nx, ny=ny, nz=nz, dh=dh, dt=dt, nt=nsteps, igreen=1,
                   src=(nx // 2, ny // 2, 10), sigma=0.3, t0=1.2, amp=1e9, nd=20,
                   npx=dec[0], npy=dec[1], npz=dec[2], ntiskp_sgt=10,
                   write_step=nsteps // 10)


def _ref_wall(d: str, dec, nsteps: int, media, pts, **grid) -> float:
    write_run(d, _spec(nsteps, dec, **grid), media, pts)
    n = dec[0] * dec[1] * dec[2]
    t = time.time()
    r = subprocess.run([MPIEXEC, "-n", str(n), REF, "IN3D"], cwd=d, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=600)
    if r.returncode != 0 or "computing time per time step" not in r.stdout:
        raise RuntimeError("reference run failed:\n" + r.stdout[-2000:])
    return time.time() - t


def measure(cores: int, short: int = 10, long: int = 40, c1: bool = True) -> Dict[str, object]:
    cores = max(1, cores)
    d = tempfile.mkdtemp(prefix="awp_cpu_")
    try:
        if os.path.exists(REF) and os.path.exists(MPIEXEC) and cores >= 4:
            dec = _decomp(cores)
            n = dec[0] * dec[1] * dec[2]
            media = layered_media(NX, NY, NZ, LAYERS)
            pts = random_cordfile_points(2000, 1, (21, NX - 20), (21, NY - 20), (1, NZ - 20))
            w1 = _ref_wall(d, dec, short, media, pts)
            w2 = _ref_wall(d, dec, long, media, pts)
            tps = (w2 - w1) / (long - short)
            out = {"value": NX * NY * NZ / tps / 1e9, "unit": "Gpt-updates/s", "cores": n,
                   "kind": "reference",
                   "sample": "reference AWP-ODC-SGT CPU solver (oracle/_ref/pmcl3d, amdflang -O2 "
                             "-fno-automatic), %dx%dx%d layered medium, h=100 m, %d MPI ranks "
                             "(%dx%dx%d); wall(%d steps) - wall(%d steps) = %.2f s -> %.4f s/step"
                             % (NX, NY, NZ, n, dec[0], dec[1], dec[2], long, short, w2 - w1, tps)}
            if c1:
                # the reference's own CPU case: 200x200x100, h = 200 m, dt = 0.01 s
                g1 = dict(nx=200, ny=200, nz=100, dh=200.0, dt=0.01)
                m1 = layered_media(200, 200, 100, LAYERS)
                p1 = random_cordfile_points(2000, 1, (21, 180), (21, 180), (1, 80))
                a = _ref_wall(d, dec, 4 * short, m1, p1, **g1)
                b = _ref_wall(d, dec, 4 * long, m1, p1, **g1)
                t1 = (b - a) / (4 * long - 4 * short)
                out["c1"] = {"value": 200 * 200 * 100 / t1 / 1e9, "unit": "Gpt-updates/s", "cores": n,
                             "sample": "200x200x100 layered (the reference's CPU case), %d MPI ranks; "
                                       "wall(%d steps) - wall(%d steps) = %.2f s -> %.5f s/step"
                                       % (n, 4 * long, 4 * short, b - a, t1)}
            return out
        # fallback: the C restatement with OpenMP
        sys.path.insert(0, REPO)
        from oracle import oracle as orc
        spec = _spec(long, (1, 1, 1))
        pts = random_cordfile_points(2000, 1, (21, NX - 20), (21, NY - 20), (1, NZ - 20))
        media = layered_media(NX, NY, NZ, LAYERS)
        write_run(d, spec, media, pts)
        p = orc.read_in3d(os.path.join(d, "IN3D"))
        xyz, stf = orc.read_source(os.path.join(d, "source"), 1, p["NST"])
        o = orc.Oracle(p, media, xyz, stf, pts)
        steps = 5
        t = time.time()
        o.run(steps, cores)
        el = time.time() - t
        return {"value": NX * NY * NZ * steps / el / 1e9, "unit": "Gpt-updates/s", "cores": cores,
                "kind": "port",
                "sample": "C restatement of the reference (oracle/awp_oracle.c, OpenMP), %dx%dx%d layered "
                          "medium, %d steps, %.1f s" % (NX, NY, NZ, steps, el)}
    finally:
        shutil.rmtree(d, ignore_errors=True)


if __name__ == "__main__":
    print(measure(int(sys.argv[1]) if len(sys.argv) > 1 else os.cpu_count()))
