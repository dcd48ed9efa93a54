"""The per-step job-value trace (CHKJ: pmcl3d.f:1476-1479, write(24,
'(I6,3E10.3)') i, u1, v1, w1 at the master rank's node (ND+2, ND+2, ND+2)).

CPU: the restatement's u1, v1, w1 at that node after every step, written in
the E10.3 edit descriptor, reproduce the reference binary's CHKJ files byte
for byte (which pins the node, the moment of the sample -- after the source
sample of the next step -- and the format).  GPU: the host executable writes
the same bytes, from single runs and from batches.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import REPO, load_golden
import cases
import awpsgt
from oracle import oracle as orc


def e10_3(v: float) -> str:
    """Fortran E10.3: 0.ddd mantissa, two-digit exponent."""
    v = float(np.float32(v))
    if v == 0.0:
        return "%10s" % ("-0.000E+00" if np.signbit(v) else "0.000E+00")
    t = "%.2e" % abs(v)
    e = int(t.split("e")[1]) + 1
    return "%s0.%s%s%sE%s%02d" % ("-" if v < 0 else " ", t[0], t[2], t[3], "-" if e < 0 else "+", abs(e))


def test_e10_3_format():
    assert e10_3(0.0) == " 0.000E+00"
    assert e10_3(-5.42e-43) == "-0.542E-42"
    assert e10_3(1.0) == " 0.100E+01"
    assert e10_3(9.9951e-5) == " 0.100E-03"
    assert e10_3(123.4) == " 0.123E+03"


@pytest.mark.parametrize("name", ["g1_fx_layered", "g5_surface_yz", "g7_random_edges"])
def test_restatement_trace_equals_reference_chkj(case_dirs, name):
    d, _ = case_dirs[name]
    q = orc.read_in3d(os.path.join(d, "IN3D"))
    media = orc.read_media(os.path.join(d, "media"), q["NX"], q["NY"], q["NZ"], q["NVAR"])
    xyz, stf = orc.read_source(os.path.join(d, "source"), 1, q["NST"])
    pts = orc.read_cordfile(os.path.join(d, "cordfile"))
    o = orc.Oracle(q, media, xyz, stf, pts)
    n = q["ND"] + 2
    lines = []
    for i in range(1, q["NT"] + 1):
        o.run(1)
        vals = [o.field(f)[n - 1, n - 1, n - 1] for f in ("u1", "v1", "w1")]
        lines.append("%6d%s%s%s\n" % (i, e10_3(vals[0]), e10_3(vals[1]), e10_3(vals[2])))
    want = load_golden(name)["chkj"].tobytes().decode()
    assert "".join(lines) == want


def _run_host(d, tmp_path, tag, args=(), env=None):
    run = os.path.join(str(tmp_path), tag)
    shutil.copytree(d, run)
    exe = os.path.join(REPO, "cybershake-core_amd", "bin", "awp_pmcl3d")
    r = subprocess.run([exe] + (list(args) or ["IN3D"]), cwd=run,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120,
                       env=dict(os.environ, **(env or {})))
    assert r.returncode == 0, r.stdout
    return run


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(cases.CASES))
def test_host_chkj_equals_reference(case_dirs, name, tmp_path):
    d, _ = case_dirs[name]
    run = _run_host(d, tmp_path, "r")
    got = open(os.path.join(run, "CHKJ"), "rb").read()
    assert got == load_golden(name)["chkj"].tobytes()


@pytest.mark.gpu
@pytest.mark.parametrize("scheme", ["split", "fused"])
def test_trace_batch_equals_single(case_dirs, scheme, monkeypatch):
    from test_gpu_batch import _make, _single
    names = ["g1_fx_layered", "g4_surface_xz", "g6_surface_zz"]
    monkeypatch.setenv("AWP_SCHEME", scheme)
    runs = [_make(case_dirs[n][0], awpsgt.AWP_STORAGE_FP32) for n in names]
    with awpsgt.Batch(runs) as b:
        b.step(runs[0].p.nt)
    for n, r in zip(names, runs):
        ref = _single(case_dirs[n][0], awpsgt.AWP_STORAGE_FP32)
        t, ijk = r.trace()
        t0, ijk0 = ref.trace()
        assert ijk == ijk0 == (10, 10, 10)
        assert t.shape == (r.p.nt, 3) and np.array_equal(t, t0), n
        assert np.any(t != 0)
        ref.close()
        r.close()
