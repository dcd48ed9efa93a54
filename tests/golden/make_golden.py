"""Generate the reference-solver fixtures under tests/golden/.

Runs the reference CPU solver (oracle/_ref/pmcl3d, built from the reference's
own Fortran sources by oracle/Makefile.ref) on the seeded cases of cases.py
with MPICH's mpiexec, and stores what it wrote:

  <case>.npz   sgt       raw SGT file of the run as float32 (whole file)
               nbytes    file size in bytes
               inputs_sha256  SHA-256 of IN3D+media+source+cordfile
               chkj      the CHKJ text file (per-step u1 v1 w1 trace), bytes
  c1_subset.npz  every 10th cordfile point of the 200x200x100 / 1500-step
               run, the SHA-256 of the whole SGT file, per-component L2
               norms over the whole file, and the solver's own timing lines.

Only runs in a container that has the reference build; the fixtures it
writes are committed and the tests read them.

    python tests/golden/make_golden.py [--c1] [--cases g1_fx_layered ...]
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import cases  # noqa: E402

REF = os.path.join(cases.REPO, "oracle", "_ref", "pmcl3d")
MPIEXEC = "/opt/conda/bin/mpiexec"


def run_reference(d: str, nranks: int, timeout: int) -> str:
    t = time.time()
    with open(os.path.join(d, "run.log"), "w") as log:
        subprocess.run([MPIEXEC, "-n", str(nranks), REF, "IN3D"], cwd=d, stdout=log,
                       stderr=subprocess.STDOUT, timeout=timeout, check=True)
    wall = time.time() - t
    with open(os.path.join(d, "run.log")) as f:
        text = f.read()
    return text + "\nWALL_SECONDS %.3f\n" % wall


def do_case(name: str) -> None:
    spec, media, pts = cases.tiny_inputs(name)
    d = tempfile.mkdtemp(prefix="awpgold_")
    try:
        sha = cases.write_case(d, spec, media, pts)
        run_reference(d, spec.npx * spec.npy * spec.npz, 600)
        raw = open(os.path.join(d, "sgt"), "rb").read()
        chkj = open(os.path.join(d, "CHKJ"), "rb").read()
        np.savez_compressed(os.path.join(HERE, name + ".npz"),
                            sgt=np.frombuffer(raw, dtype=np.float32),
                            nbytes=np.int64(len(raw)), inputs_sha256=np.array(sha),
                            chkj=np.frombuffer(chkj, dtype=np.uint8))
        print(name, "bytes", len(raw), "sha", sha[:12])
    finally:
        shutil.rmtree(d, ignore_errors=True)


def do_c1() -> None:
    spec, media, pts = cases.c1_inputs()
    d = tempfile.mkdtemp(prefix="awpgold_c1_")
    try:
        sha = cases.write_case(d, spec, media, pts)
        log = run_reference(d, spec.npx * spec.npy * spec.npz, 7200)
        raw = open(os.path.join(d, "sgt"), "rb").read()
        a = np.frombuffer(raw, dtype=np.float32)
        nt_out = 150
        a3 = a.reshape(len(pts), 6, nt_out)
        sub = np.arange(0, len(pts), 10)
        timing = [ln.strip() for ln in log.splitlines()
                  if re.search(r"time|WALL_SECONDS", ln, re.I) and "TIME STEP" not in ln]
        np.savez_compressed(os.path.join(HERE, "c1_subset.npz"),
                            idx=sub.astype(np.int32), sgt=a3[sub].copy(),
                            file_sha256=np.array(hashlib.sha256(raw).hexdigest()),
                            nbytes=np.int64(len(raw)),
                            comp_l2=np.array([np.linalg.norm(a3[:, c].astype(np.float64))
                                              for c in range(6)]),
                            inputs_sha256=np.array(sha),
                            timing=np.array(json.dumps(timing)))
        print("c1 bytes", len(raw), "timing", timing[-6:])
    finally:
        shutil.rmtree(d, ignore_errors=True)


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--c1", action="store_true")
    ap.add_argument("--cases", nargs="*", default=None)
    a = ap.parse_args()
    if not os.path.exists(REF):
        sys.exit("reference build missing: make -f oracle/Makefile.ref")
    names = a.cases if a.cases is not None else list(cases.CASES)
    for n in names:
        do_case(n)
    if a.c1:
        do_c1()


if __name__ == "__main__":
    main()
