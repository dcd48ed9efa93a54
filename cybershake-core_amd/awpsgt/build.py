"""Build libawpsgt.so (HIP kernels + C ABI + C host formats) for gfx950 and the
C host executable awp_pmcl3d, in-tree under cybershake-core_amd/{lib,bin}.

    python -m awpsgt.build        (from cybershake-core_amd/)

Device code is compiled with -ffp-contract=off so every FP32 statement is the
IEEE result of the reference's own operation sequence, and without packed FP32
(-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REPO = os.path.dirname(PKG)
INCLUDE = os.path.join(REPO, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value", "-I" + INCLUDE,
]
C_FLAGS = ["-std=c11", "-O2", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter", "-I" + INCLUDE]

HIP_SOURCES = ["csrc/awp_kernels.hip", "csrc/awp_fused.hip", "csrc/awp_ctx.hip", "csrc/awp_batch.hip"]
C_SOURCES = ["host/awp_io.c", "host/awp_head.c", "host/awp_sgtcheck.c"]


def _run(cmd, quiet=True):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(ln for ln in r.stdout.splitlines()
                    if "packed-fp32-ops' is not a recognized feature" not in ln)
    if r.returncode != 0:
        raise RuntimeError("build step failed (%d): %s\n%s" % (r.returncode, " ".join(cmd), out))
    if out.strip() and not quiet:
        print(out)


def build(verbose: bool = False) -> str:
    """Compile everything; returns the path of libawpsgt.so."""
    libdir = os.path.join(PKG, "lib")
    bindir = os.path.join(PKG, "bin")
    os.makedirs(libdir, exist_ok=True)
    os.makedirs(bindir, exist_ok=True)
    tmp = tempfile.mkdtemp(prefix="awpsgt_build_")
    try:
        objs = []
        for src in HIP_SOURCES:
            o = os.path.join(tmp, os.path.basename(src) + ".o")
            _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(PKG, src), "-o", o], quiet=not verbose)
            objs.append(o)
        for src in C_SOURCES:
            o = os.path.join(tmp, os.path.basename(src) + ".o")
            _run(["gcc"] + C_FLAGS + ["-c", os.path.join(PKG, src), "-o", o], quiet=not verbose)
            objs.append(o)
        so_tmp = os.path.join(tmp, "libawpsgt.so")
        _run([HIPCC, "-shared", "--offload-arch=" + ARCH, "-o", so_tmp] + objs, quiet=not verbose)
        so = os.path.join(libdir, "libawpsgt.so")
        shutil.move(so_tmp, so)
        exe_tmp = os.path.join(tmp, "awp_pmcl3d")
        _run(["gcc"] + C_FLAGS + [os.path.join(PKG, "host", "awp_pmcl3d.c"), "-o", exe_tmp,
                                  "-L" + libdir, "-lawpsgt", "-Wl,-rpath,$ORIGIN/../lib", "-lm"],
             quiet=not verbose)
        shutil.move(exe_tmp, os.path.join(bindir, "awp_pmcl3d"))
        tool_tmp = os.path.join(tmp, "awp_sgt")
        _run(["gcc"] + C_FLAGS + [os.path.join(PKG, "host", "awp_sgt.c"), "-o", tool_tmp,
                                  "-L" + libdir, "-lawpsgt", "-Wl,-rpath,$ORIGIN/../lib", "-lm"],
             quiet=not verbose)
        shutil.move(tool_tmp, os.path.join(bindir, "awp_sgt"))
        return so
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
