"""Checkpoint / restart through the C ABI and the C host.

A run saved after k steps and restored into a fresh run of the same set-up
must continue bit for bit as the uninterrupted run: same fields after the
last step, same SGT samples, same SGT file -- in both storage modes (the
FP16 memory variables and their exponents are part of the saved state) and
both time-step schemes (the fused scheme's buffer parity too).  The host
(awp_pmcl3d with CHECKPOINT > 0 in the IN3D) writes output_ckp/awp_ckp.bin
every CHECKPOINT steps and, started again, resumes from it.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import awpsgt
from test_gpu_batch import _make, _single

pytestmark = pytest.mark.gpu

CASE = "g7_random_edges"   # 133 steps, points on every edge


@pytest.mark.parametrize("scheme", ["split", "fused"])
@pytest.mark.parametrize("storage", [awpsgt.AWP_STORAGE_FP32, awpsgt.AWP_STORAGE_MIXED])
def test_restart_continues_bit_for_bit(case_dirs, storage, scheme, tmp_path, monkeypatch):
    d = case_dirs[CASE][0]
    monkeypatch.setenv("AWP_SCHEME", scheme)
    ref = _single(d, storage)            # uninterrupted
    a = _make(d, storage)
    a.step(37)                           # odd: the fused scheme's buffer parity is 1
    ckp = os.path.join(str(tmp_path), "ckp.bin")
    a.checkpoint_save(ckp)
    assert not os.path.exists(ckp + ".tmp")
    a.step(a.p.nt)
    a.synchronize()
    b = _make(d, storage)
    b.checkpoint_load(ckp)
    assert b.progress()[0] == 37
    b.step(b.p.nt)
    b.synchronize()
    for r in (a, b):
        assert r.progress()[0] == ref.p.nt
        assert np.array_equal(r.sgt(), ref.sgt())
        for f in ("u1", "w1", "xx", "xz", "r1", "r5"):
            k = 2 if f in ("u1", "w1", "xz") else 0
            assert np.array_equal(r.field(f, k), ref.field(f, k)), f
    fa, fb = os.path.join(str(tmp_path), "fa"), os.path.join(str(tmp_path), "fb")
    ref.write_sgt_file(fa)
    b.write_sgt_file(fb)
    assert open(fa, "rb").read() == open(fb, "rb").read()
    for r in (ref, a, b):
        r.close()


def test_restart_rejects_other_runs(case_dirs, tmp_path):
    d = case_dirs[CASE][0]
    a = _make(d, awpsgt.AWP_STORAGE_FP32)
    a.step(5)
    ckp = os.path.join(str(tmp_path), "ckp.bin")
    a.checkpoint_save(ckp)
    m = _make(d, awpsgt.AWP_STORAGE_MIXED)
    with pytest.raises(awpsgt.AwpError, match="another run"):
        m.checkpoint_load(ckp)
    other = _make(case_dirs["g1_fx_layered"][0], awpsgt.AWP_STORAGE_FP32)
    with pytest.raises(awpsgt.AwpError, match="another run"):
        other.checkpoint_load(ckp)
    bad = os.path.join(str(tmp_path), "bad.bin")
    open(bad, "wb").write(b"not a checkpoint" * 8)
    with pytest.raises(awpsgt.AwpError, match="not a checkpoint"):
        a.checkpoint_load(bad)
    for r in (a, m, other):
        r.close()


def test_host_checkpoint_and_resume(case_dirs, tmp_path):
    """CHECKPOINT=40 in the IN3D: the first host run writes the checkpoint
    at steps 40, 80, 120; a second run in the same directory resumes at 120
    and writes the same SGT file as the uninterrupted library run."""
    from conftest import REPO
    exe = os.path.join(REPO, "cybershake-core_amd", "bin", "awp_pmcl3d")
    run = os.path.join(str(tmp_path), "run")
    shutil.copytree(case_dirs[CASE][0], run)
    os.makedirs(os.path.join(run, "output_ckp"))
    lines = open(os.path.join(run, "IN3D")).read().splitlines()
    lines = [("       40 CHECKPOINT" if ln.split()[1:2] == ["CHECKPOINT"] else ln) for ln in lines]
    open(os.path.join(run, "IN3D"), "w").write("\n".join(lines) + "\n")
    r1 = subprocess.run([exe, "IN3D"], cwd=run, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                        timeout=300)
    assert r1.returncode == 0, r1.stdout
    assert "Writing checkpoint file at step 120" in r1.stdout
    first = open(os.path.join(run, "sgt"), "rb").read()
    os.remove(os.path.join(run, "sgt"))
    r2 = subprocess.run([exe, "IN3D"], cwd=run, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                        timeout=300)
    assert r2.returncode == 0, r2.stdout
    assert "Checkpoint found, starting from step 120" in r2.stdout
    second = open(os.path.join(run, "sgt"), "rb").read()
    ref = _single(case_dirs[CASE][0], awpsgt.AWP_STORAGE_FP32)
    want = os.path.join(str(tmp_path), "want")
    ref.write_sgt_file(want)
    ref.close()
    assert first == second == open(want, "rb").read()
