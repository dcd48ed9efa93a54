"""FP64 reference SGT for the mixed-storage error statement.

Runs the FP64 build of the CPU restatement (oracle/awp_oracle.c, REAL=double)
on the 800x600x200 / 1500-step case of cases.C4 and stores every 4th of its
2000 SGT points (rounded to float32: 6e-8 relative, far below the errors it
measures) with the per-component L2 norms over all points:

    tests/golden/c4_fp64.npz   idx, sgt [500][6][150], comp_l2 [6], inputs_sha256

About 1.5 h on 8 cores.  (The FP32 error is measured on the GPU, whose
FP32 path is bit-identical to the FP32 restatement.)

    python tests/golden/make_fp64.py [threads]
"""
import os
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import cases  # noqa: E402

sys.path.insert(0, cases.REPO)
from oracle import oracle as orc  # noqa: E402


def main():
    nthreads = int(sys.argv[1]) if len(sys.argv) > 1 else os.cpu_count()
    spec, media, pts = cases.c4_inputs()
    d = tempfile.mkdtemp(prefix="awp_c4_")
    sha = cases.write_case(d, spec, media, pts)
    p = orc.read_in3d(os.path.join(d, "IN3D"))
    xyz, stf = orc.read_source(os.path.join(d, "source"), p["NSRC"], p["NST"])
    out = {}
    for dt in (np.float64,):
        o = orc.Oracle(p, media, xyz, stf, pts, dtype=dt)
        t = time.time()
        for s in range(0, p["NT"], 100):
            o.run(min(100, p["NT"] - s), nthreads)
            print(dt.__name__, "step", o.info()["step"], "%.0f s" % (time.time() - t), flush=True)
        out[dt] = o.sgt()[0]
        o.close()
    s64 = out[np.float64]
    idx = np.arange(0, len(pts), 4)
    comp_l2 = np.array([np.linalg.norm(s64[:, c]) for c in range(6)])
    np.savez_compressed(os.path.join(HERE, "c4_fp64.npz"), idx=idx.astype(np.int32),
                        sgt=s64[idx].astype(np.float32), comp_l2=comp_l2,
                        inputs_sha256=np.array(sha))
    print("done; comp_l2", comp_l2)


if __name__ == "__main__":
    main()
