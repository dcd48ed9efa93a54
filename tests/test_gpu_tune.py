"""The single-run tuner and scheme switches (awp_tune, awp_set_scheme,
awp_get_scheme) through the C ABI.  Whatever launch shape the tuner keeps,
and wherever a run switches between the fused and the two-stage step (also
while the fused step's state sits in its second wavefield buffer), the run's
SGT samples and fields equal those of the same run stepped two-stage
throughout, bit for bit; a checkpoint written by the fused step loads into
a run on the two-stage step and continues the same way."""
import os

import numpy as np
import pytest

import awpsgt
from test_gpu_batch import _make, _single

pytestmark = pytest.mark.gpu

CASE = "g7_random_edges"   # 133 steps, points on every edge
STORAGES = [awpsgt.AWP_STORAGE_FP32, awpsgt.AWP_STORAGE_MIXED]


def _two_stage(d, storage, monkeypatch):
    monkeypatch.setenv("AWP_SCHEME", "split")
    r = _single(d, storage)
    monkeypatch.delenv("AWP_SCHEME")
    assert not r.scheme()[0]
    return r


def _same(r, ref):
    assert r.progress() == ref.progress()
    assert np.array_equal(r.sgt(), ref.sgt())
    for f in ("u1", "v1", "w1", "xx", "yy", "zz", "xy", "xz", "yz", "r1", "r5"):
        k = 2 if f in ("u1", "v1", "w1", "zz", "xz", "yz") else 0
        assert np.array_equal(r.field(f, k), ref.field(f, k)), f


@pytest.mark.parametrize("storage", STORAGES)
def test_tuned_run_equals_two_stage(case_dirs, storage, monkeypatch):
    d = case_dirs[CASE][0]
    ref = _two_stage(d, storage, monkeypatch)
    monkeypatch.setenv("AWP_SCHEME", "fused")    # both buffers: every candidate is open
    r = _make(d, storage)
    fused, ty, nc = r.tune(2)
    assert r.progress() == (0, 0)
    assert (fused and ty in (4, 7, 8, 11)) or (not fused and ty in (1, 2, 4, 8, 16))
    assert nc >= 1
    r.step(r.p.nt)
    r.synchronize()
    _same(r, ref)
    for x in (r, ref):
        x.close()


@pytest.mark.parametrize("storage", STORAGES)
def test_switch_schemes_between_steps(case_dirs, storage, monkeypatch):
    d = case_dirs[CASE][0]
    ref = _two_stage(d, storage, monkeypatch)
    monkeypatch.setenv("AWP_SCHEME", "fused")
    r = _make(d, storage)
    assert r.scheme()[0]
    r.step(37)                 # odd: the state sits in the second buffer
    r.set_scheme(False)
    assert not r.scheme()[0]
    r.step(20)
    r.set_scheme(True)
    r.step(15)                 # odd again
    r.set_scheme(False)
    r.step(r.p.nt)
    r.synchronize()
    _same(r, ref)
    for x in (r, ref):
        x.close()


def test_fused_checkpoint_loads_into_two_stage(case_dirs, tmp_path, monkeypatch):
    d = case_dirs[CASE][0]
    ref = _two_stage(d, awpsgt.AWP_STORAGE_FP32, monkeypatch)
    monkeypatch.setenv("AWP_SCHEME", "fused")
    a = _make(d, awpsgt.AWP_STORAGE_FP32)
    a.step(37)
    ckp = os.path.join(str(tmp_path), "ckp.bin")
    a.checkpoint_save(ckp)
    b = _make(d, awpsgt.AWP_STORAGE_FP32)
    b.set_scheme(False)
    b.checkpoint_load(ckp)
    assert b.progress()[0] == 37 and not b.scheme()[0]
    b.step(b.p.nt)
    b.synchronize()
    _same(b, ref)
    for x in (a, b, ref):
        x.close()


def test_scheme_errors(case_dirs, monkeypatch):
    d = case_dirs[CASE][0]
    monkeypatch.setenv("AWP_SCHEME", "split")
    r = _make(d, awpsgt.AWP_STORAGE_FP32)
    with pytest.raises(awpsgt.AwpError, match="second wavefield buffer"):
        r.set_scheme(True)
    r.step(3)
    with pytest.raises(awpsgt.AwpError, match="not stepped"):
        r.tune()
    r.close()


@pytest.mark.parametrize("nz", [33, 34])
def test_fused_pairs_meet_the_free_surface(nz, tmp_path, monkeypatch):
    """Two levels per barrier on 64 x 11 tiles: with nz = 33 the level pair
    (nz, nz + 1) falls in one group, so the free-surface images at nz + 1
    wait for level nz behind an extra barrier; with nz = 34 they do not.
    Both equal the two-stage step bit for bit, in one z chunk and in three."""
    import cases
    spec = cases.tiny_spec("g1_fx_layered")
    spec.nx, spec.ny, spec.nz = 70, 30, nz
    spec.npx = spec.npy = spec.npz = 1
    spec.src = (35, 15, 6)
    media = awpsgt.inputs.layered_media(70, 30, nz, [700.0, 1500.0, 2500.0])
    pts = awpsgt.inputs.random_cordfile_points(60, 5, (1, 70), (1, 30), (1, nz))
    d = str(tmp_path)
    cases.write_case(d, spec, media, pts)
    ref = _two_stage(d, awpsgt.AWP_STORAGE_FP32, monkeypatch)
    monkeypatch.setenv("AWP_SCHEME", "fused")
    monkeypatch.setenv("AWP_TILE_Y", "11")
    monkeypatch.setenv("AWP_PF", "3")
    for kc in ("40", "12"):
        monkeypatch.setenv("AWP_KC", kc)
        r = _make(d, awpsgt.AWP_STORAGE_FP32)
        assert r.scheme() == (True, 11, 1 if kc == "40" else 3)
        r.step(r.p.nt)
        r.synchronize()
        _same(r, ref)
        r.close()
    ref.close()
