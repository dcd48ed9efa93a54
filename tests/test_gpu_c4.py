"""SGT error of the device runs against the FP64 run of the CPU restatement on
the 800x600x200 / 1500-step case (layered medium of the 200x200x100
reference case, 2000 SGT points; tests/golden/make_fp64.py), per component:
relative L2 over every 4th point and peak error relative to the component's
peak.  Bounds: FP32 <= 1e-4 (the FP32 parity bound), mixed storage (FP16
r1..r6, FP16 qp/qs) <= 1e-3, and the mixed run's FP16 SGT file (per point and
component FP32 scale) <= 1e-3 as well."""
import os

import numpy as np
import pytest

from conftest import load_golden, rel_l2
import awpsgt
from test_gpu_parity import _inputs

pytestmark = pytest.mark.gpu

COMP = ("exx", "eyy", "ezz", "exy", "exz", "eyz")


def _errors(a, ref):
    rel = [rel_l2(a[:, c], ref[:, c]) for c in range(6)]
    peak = [float(np.max(np.abs(a[:, c].astype(np.float64) - ref[:, c])) / np.max(np.abs(ref[:, c])))
            for c in range(6)]
    return rel, peak


def test_c4_sgt_error_vs_fp64(c4_dir, tmp_path):
    d, sha = c4_dir
    g = load_golden("c4_fp64")
    assert str(g["inputs_sha256"]) == sha
    idx, ref = g["idx"], g["sgt"].astype(np.float64)
    p, media, xyz, stf, pts = _inputs(d)
    rows = []
    for storage, tol in ((awpsgt.AWP_STORAGE_FP32, 1e-4), (awpsgt.AWP_STORAGE_MIXED, 1e-3)):
        r = awpsgt.Run(p, storage=storage)
        r.set_sgt_points(pts)
        r.set_media(media)
        r.set_source(xyz, stf)
        r.init()
        r.step(p.nt)
        r.synchronize()
        s = r.sgt()[idx]
        rel, peak = _errors(s, ref)
        name = "fp32" if storage == awpsgt.AWP_STORAGE_FP32 else "mixed"
        rows.append((name, rel, peak))
        assert max(rel) <= tol, (name, rel)
        if storage == awpsgt.AWP_STORAGE_MIXED:
            f16 = os.path.join(str(tmp_path), "sgt16")
            r.write_sgt_file_fp16(f16)
            dec = awpsgt.read_sgt16(f16)[idx]
            rel16, peak16 = _errors(dec, ref)
            rows.append(("mixed+fp16 file", rel16, peak16))
            assert max(rel16) <= tol, rel16
        r.close()
    for name, rel, peak in rows:
        print("%-16s rel L2 %s | peak %s" % (name, " ".join("%s %.2e" % (c, e) for c, e in zip(COMP, rel)),
                                               " ".join("%.2e" % e for e in peak)))
