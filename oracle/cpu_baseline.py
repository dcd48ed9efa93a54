"""CPU baseline timing for bench.py (test infrastructure: the cpu_baseline
leg is the only product-side caller, and only to time the checker).

Preferred: the reference CPU solver itself (oracle/_ref/pmcl3d, built from
the reference Fortran by oracle/Makefile.ref) run with MPICH's mpiexec on the
host cores on the 200x200x100 layered case, for a bounded number of steps; its
own "computing time per time step" line (pmcl3d.f:2151) is the measurement.
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


def _decomp(cores: int):
    """Largest NPX x NPY x NPZ <= cores with >= 2 ranks per horizontal axis,
    each rank at least 20 cells wide on the 200x200x100 grid."""
    best = (2, 2, 1)
    for px in (2, 4, 5):
        for py in (2, 4, 5):
            for pz in (1, 2, 4, 5):
                if px * py * pz <= cores and px * py * pz > best[0] * best[1] * best[2]:
                    best = (px, py, pz)
    return best


def _spec(nsteps: int, dec) -> RunSpec:
    return RunSpec(nx=200, ny=200, nz=100, dh=200.0, dt=0.01, nt=nsteps, igreen=1,
                   src=(100, 100, 10), sigma=0.3, t0=1.2, amp=1e9, nd=20,
                   npx=dec[0], npy=dec[1], npz=dec[2], ntiskp_sgt=10,
                   write_step=nsteps // 10)


def measure(cores: int, nsteps: int = 300) -> Dict[str, object]:
    cores = max(1, cores)
    d = tempfile.mkdtemp(prefix="awp_cpu_")
    try:
        if os.path.exists(REF) and os.path.exists(MPIEXEC) and cores >= 4:
            dec = _decomp(cores)
            spec = _spec(nsteps, dec)
            pts = random_cordfile_points(2000, 1, (21, 180), (21, 180), (1, 80))
            write_run(d, spec, layered_media(200, 200, 100, LAYERS), pts)
            n = dec[0] * dec[1] * dec[2]
            t = time.time()
            r = subprocess.run([MPIEXEC, "-n", str(n), REF, "IN3D"], cwd=d, stdout=subprocess.PIPE,
                               stderr=subprocess.STDOUT, text=True, timeout=600)
            wall = time.time() - t
            m = re.search(r"computing time per time step\s*=\s*([0-9.Ee+-]+)", r.stdout)
            if r.returncode == 0 and m:
                tps = float(m.group(1))
                return {"value": 4.0e6 / tps / 1e9, "unit": "Gpt-updates/s", "cores": n,
                        "kind": "reference",
                        "sample": "reference AWP-ODC-SGT CPU solver (oracle/_ref/pmcl3d, amdflang -O2), "
                                  "200x200x100 layered case, %d steps, %d MPI ranks (%dx%dx%d); "
                                  "its own 'computing time per time step' = %.4f s; wall %.1f s"
                                  % (nsteps, n, dec[0], dec[1], dec[2], tps, wall)}
        # fallback: the C restatement with OpenMP
        sys.path.insert(0, REPO)
        from oracle import oracle as orc
        spec = _spec(nsteps, (1, 1, 1))
        pts = random_cordfile_points(2000, 1, (21, 180), (21, 180), (1, 80))
        media = layered_media(200, 200, 100, LAYERS)
        write_run(d, spec, media, pts)
        p = orc.read_in3d(os.path.join(d, "IN3D"))
        xyz, stf = orc.read_source(os.path.join(d, "source"), 1, p["NST"])
        o = orc.Oracle(p, media, xyz, stf, pts)
        steps = min(nsteps, 100)
        t = time.time()
        o.run(steps, cores)
        el = time.time() - t
        return {"value": 4.0e6 * steps / el / 1e9, "unit": "Gpt-updates/s", "cores": cores, "kind": "port",
                "sample": "C restatement of the reference (oracle/awp_oracle.c, OpenMP), 200x200x100 layered "
                          "case, %d steps, %.1f s" % (steps, el)}
    finally:
        shutil.rmtree(d, ignore_errors=True)


if __name__ == "__main__":
    print(measure(int(sys.argv[1]) if len(sys.argv) > 1 else os.cpu_count()))
