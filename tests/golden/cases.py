"""Seeded SGT run cases shared by the fixture generator (make_golden.py) and
the tests.  Each case is fully determined by the parameters below and the
generators in awpsgt.inputs, so inputs are regenerated on demand and only the
reference solver's outputs are stored as fixtures."""
from __future__ import annotations

import hashlib
import os
import sys
from typing import Dict

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
PKG = os.path.join(REPO, "cybershake-core_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)

from awpsgt.inputs import (RunSpec, layered_media, random_media,  # noqa: E402
                           random_cordfile_points, write_run)

LAYERS_C1 = [600.0, 1400.0, 2200.0, 3000.0, 3800.0]

# name -> dict of case options
CASES: Dict[str, Dict] = {
    # igreen 1 body force, layered medium, 2x2x2 ranks
    "g1_fx_layered": dict(igreen=1, src=(20, 17, 5)),
    # igreen 2, explicit Q columns (NVAR=5), no dispersion correction
    "g2_fy_nvar5_nosocalq": dict(igreen=2, src=(30, 25, 7), nvar=5, socalq=0),
    # igreen 3 body force at the free surface
    "g3_fz_surface": dict(igreen=3, src=(24, 20, 1)),
    # igreen 4 surface traction source (xz), the CPU builder's mode
    "g4_surface_xz": dict(igreen=4, src=(24, 20, 1)),
    # igreen 5 surface traction source (yz), 2x4x2 ranks, NVAR 5 with
    # SOCALQ off
    "g5_surface_yz": dict(igreen=5, src=(17, 26, 1), npx=2, npy=4, npz=2, nvar=5, socalq=0),
    # igreen 6 surface normal-stress source (zz)
    "g6_surface_zz": dict(igreen=6, src=(30, 12, 1)),
    # random medium, 4x2x2 ranks, WRITE_STEP 5 (two flushes, a dropped partial
    # buffer, holes in the file) and edge-case cordfile points: beyond the far
    # x/y faces and above the surface (dropped), x=0 / y=0 / z=nz+1 (the
    # reference maps them onto a node) and the grid corners.
    "g7_random_edges": dict(igreen=1, src=(22, 19, 9), rand=True, ws=5, nt=133,
                            npx=4, npy=2, npz=2,
                            extra_pts=[[49, 3, 3], [0, 5, 5], [5, 0, 5], [5, 5, 33],
                                       [5, 5, 0], [48, 40, 32], [1, 1, 1], [7, 41, 2]]),
}

C1 = dict(nx=200, ny=200, nz=100, dh=200.0, dt=0.01, nt=1500, igreen=1, src=(100, 100, 10),
          sigma=0.3, t0=1.2, amp=1e9, nd=20, npx=2, npy=2, npz=2, ntiskp_sgt=10, ws=150,
          npts=2000, pts_seed=1, pts_range=((21, 180), (21, 180), (1, 80)))


# Mixed-storage error case: the 200x200x100 recipe (layers, h, dt, steps) on
# 800x600x200; the FP64 restatement is its reference.
C4 = dict(nx=800, ny=600, nz=200, dh=200.0, dt=0.01, nt=1500, igreen=1, src=(400, 300, 10),
          sigma=0.3, t0=1.2, amp=1e9, nd=20, ntiskp_sgt=10, ws=150,
          npts=2000, pts_seed=4, pts_range=((21, 780), (21, 580), (1, 160)))


def c4_spec() -> RunSpec:
    c = C4
    return RunSpec(nx=c["nx"], ny=c["ny"], nz=c["nz"], dh=c["dh"], dt=c["dt"], nt=c["nt"],
                   igreen=c["igreen"], src=c["src"], sigma=c["sigma"], t0=c["t0"],
                   amp=c["amp"], nd=c["nd"], ntiskp_sgt=c["ntiskp_sgt"], write_step=c["ws"])


def c4_inputs():
    spec = c4_spec()
    media = layered_media(spec.nx, spec.ny, spec.nz, LAYERS_C1)
    (xr, yr, zr) = C4["pts_range"]
    pts = random_cordfile_points(C4["npts"], C4["pts_seed"], xr, yr, zr)
    return spec, media, pts


def tiny_spec(name: str) -> RunSpec:
    c = CASES[name]
    return RunSpec(nx=48, ny=40, nz=32, dh=200.0, dt=0.01, nt=c.get("nt", 120),
                   igreen=c["igreen"], src=c["src"], nd=8, npx=c.get("npx", 2),
                   npy=c.get("npy", 2), npz=c.get("npz", 2), ntiskp_sgt=10, sigma=0.15,
                   t0=0.4, nvar=c.get("nvar", 3), socalq=c.get("socalq", 1),
                   write_step=c.get("ws"))


def tiny_inputs(name: str):
    c = CASES[name]
    spec = tiny_spec(name)
    nx, ny, nz = spec.nx, spec.ny, spec.nz
    if c.get("rand"):
        media = random_media(nx, ny, nz, seed=5, vs_top=800.0, vs_bot=3500.0, pert=0.1)
    else:
        media = layered_media(nx, ny, nz, LAYERS_C1)
    if spec.nvar == 5:
        vs = media[..., 1]
        media = np.stack([media[..., 0], vs, media[..., 2], 0.1 * vs, 0.05 * vs],
                         axis=-1).astype(np.float32)
    pts = random_cordfile_points(60, 1, (1, nx), (1, ny), (1, nz))
    if c.get("extra_pts"):
        pts = np.concatenate([pts, np.array(c["extra_pts"], dtype=np.int64)])
    return spec, media, pts


def c1_spec() -> RunSpec:
    c = C1
    return RunSpec(nx=c["nx"], ny=c["ny"], nz=c["nz"], dh=c["dh"], dt=c["dt"], nt=c["nt"],
                   igreen=c["igreen"], src=c["src"], sigma=c["sigma"], t0=c["t0"],
                   amp=c["amp"], nd=c["nd"], npx=c["npx"], npy=c["npy"], npz=c["npz"],
                   ntiskp_sgt=c["ntiskp_sgt"], write_step=c["ws"])


def c1_inputs():
    spec = c1_spec()
    media = layered_media(spec.nx, spec.ny, spec.nz, LAYERS_C1)
    (xr, yr, zr) = C1["pts_range"]
    pts = random_cordfile_points(C1["npts"], C1["pts_seed"], xr, yr, zr)
    return spec, media, pts


def write_case(dirpath: str, spec, media, pts) -> str:
    """Writes the run directory and returns the SHA-256 of its four inputs."""
    write_run(dirpath, spec, media, pts)
    h = hashlib.sha256()
    for f in ("IN3D", "media", "source", "cordfile"):
        with open(os.path.join(dirpath, f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()
