"""Writes tests/golden/sgt_head.npz: the header files the reference's SgtHead
write_head (built from its sources into oracle/_ref/write_head) makes for the
seeded case of head_case.py, with and without its -c (corrected mu) flag,
for the x and z components."""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import head_case as hc  # noqa: E402

TOOL = os.path.join(HERE, "..", "..", "oracle", "_ref", "write_head")


def run_tool(p: dict, comp: str, corr: bool, out: str) -> bytes:
    a = hc.ARGS
    cmd = [TOOL, p["modelbox"], p["coordfile"], p["fdloc"], p["gridout"], str(a["spacing"]), str(a["nt"]),
           str(a["dt"]), str(a["sgt_tinc"]), comp, "%g" % a["moment"], str(a["flo"]), p["media"], out]
    if corr:
        cmd.append("-c")
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    return open(out, "rb").read()


def main():
    d = tempfile.mkdtemp()
    p = hc.write_inputs(d)
    out = {}
    for comp in ("x", "z"):
        for corr in (False, True):
            key = "%s_%s" % (comp, "corr" if corr else "plain")
            out[key] = np.frombuffer(run_tool(p, comp, corr, os.path.join(d, key)), dtype=np.uint8)
    np.savez_compressed(os.path.join(HERE, "sgt_head.npz"), **out)
    print({k: v.size for k, v in out.items()})


if __name__ == "__main__":
    main()
