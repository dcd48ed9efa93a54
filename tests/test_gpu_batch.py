"""Batched runs (awp_batch_*) through the C ABI.

A batch advances several runs with one persistent launch per stage of a
step over all their tiles (two-stage scheme: velocity, free surface, stress;
fused scheme: one sweep) and replays the steps from a HIP graph.  Every
run's SGT and fields must equal, bit for bit, those of the same run stepped
on its own (which the single-run tests tie to the oracle and the reference
files), whatever the step counts per call (graph pairs plus an odd
remainder), the scheme and kernel variant the tuner picks, and the storage
mode.
"""
import os

import numpy as np
import pytest

import cases
import awpsgt
from test_gpu_parity import _inputs

pytestmark = pytest.mark.gpu

SAME_NT = ["g1_fx_layered", "g2_fy_nvar5_nosocalq", "g3_fz_surface", "g4_surface_xz", "g5_surface_yz",
           "g6_surface_zz"]


def _make(d, storage, init=True):
    p, media, xyz, stf, pts = _inputs(d)
    r = awpsgt.Run(p, storage=storage)
    r.set_media(media)
    r.set_source(xyz, stf)
    r.set_sgt_points(pts)
    if init:
        r.init()
    return r


def _single(d, storage, nsteps=None):
    r = _make(d, storage)
    r.step(r.p.nt if nsteps is None else nsteps)
    r.synchronize()
    return r


@pytest.mark.parametrize("scheme", ["split", "fused", "split_ty16", "split_xw2", "split_o32v", "fused_ty11",
                                    "split_nograph", "fused_nograph"])
@pytest.mark.parametrize("storage", [awpsgt.AWP_STORAGE_FP32, awpsgt.AWP_STORAGE_MIXED])
def test_batch_equals_single_runs(case_dirs, storage, scheme, monkeypatch):
    dirs = [case_dirs[n][0] for n in SAME_NT]
    monkeypatch.setenv("AWP_SCHEME", "fused" if scheme.startswith("fused") else "split")
    if scheme == "fused_ty11":
        monkeypatch.setenv("AWP_TILE_Y", "11")
    if scheme == "split_ty16":
        monkeypatch.setenv("AWP_TILE_Y", "16")
    if scheme == "split_xw2":
        monkeypatch.setenv("AWP_TILE_XW", "2")
    if scheme == "split_o32v":
        monkeypatch.setenv("AWP_O32V", "1")
    if scheme.endswith("_nograph"):
        # each step's launches issued directly instead of replaying the graph
        monkeypatch.setenv("AWP_GRAPH", "0")
    runs = [_make(d, storage) for d in dirs]
    monkeypatch.delenv("AWP_SCHEME", raising=False)
    nt = runs[0].p.nt
    assert all(r.p.nt == nt for r in runs)
    with awpsgt.Batch(runs) as b:
        assert b.fused() == scheme.startswith("fused")
        if scheme == "split_ty16":
            assert b.launch()[0] == 16
        if scheme == "fused_ty11":
            assert b.launch()[0] == 11
        # uneven calls: graph pairs, odd remainders, a call past the end
        for n in (1, 36, 7, 40, 1000):
            b.step(n)
    for r in runs:
        assert r.progress()[0] == nt
    for name, d, r in zip(SAME_NT, dirs, runs):
        ref = _single(d, storage)
        assert np.array_equal(r.sgt(), ref.sgt()), name
        for f in ("u1", "w1", "xx", "xz", "r1"):
            k = 2 if f in ("u1", "w1", "xz") else 0
            assert np.array_equal(r.field(f, k), ref.field(f, k)), (name, f)
        ref.close()
    for r in runs:
        r.close()


@pytest.mark.parametrize("scheme", ["split", "fused"])
def test_batch_uneven_grids_and_tuner(case_dirs, tmp_path, scheme, monkeypatch):
    """Runs of different grid shapes in one batch (their tiles interleave in
    the persistent grids), after the tuner has picked a variant (among both
    schemes when the runs hold the fused scheme's buffers); files equal the
    single-run files byte for byte."""
    import tempfile
    dirs = [case_dirs["g1_fx_layered"][0]]
    # two more shapes with the same step count
    for k, (nx, ny, nz) in enumerate([(70, 34, 40), (40, 58, 24)]):
        spec = cases.tiny_spec("g1_fx_layered")
        spec.nx, spec.ny, spec.nz = nx, ny, nz
        spec.src = (nx // 2, ny // 2, 6)
        media = awpsgt.inputs.layered_media(nx, ny, nz, [700.0, 1500.0, 2500.0])
        pts = awpsgt.inputs.random_cordfile_points(50, 11 + k, (1, nx), (1, ny), (1, nz))
        d = tempfile.mkdtemp()
        cases.write_case(d, spec, media, pts)
        dirs.append(d)
    monkeypatch.setenv("AWP_SCHEME", scheme)
    runs = [_make(d, awpsgt.AWP_STORAGE_FP32) for d in dirs]
    monkeypatch.delenv("AWP_SCHEME", raising=False)
    nt = min(r.p.nt for r in runs)
    with awpsgt.Batch(runs) as b:
        ty, pf, nc = b.tune(1)
        if b.fused():
            assert scheme == "fused"
            assert ty in (4, 7, 8, 11) and pf in (0, 1, 2, 3) and nc >= 1
        else:
            assert ty in (4, 8, 16) and pf in (1, 2) and nc >= 1
        assert all(r.progress() == (0, 0) for r in runs)
        b.step(nt)
    for k, (d, r) in enumerate(zip(dirs, runs)):
        ref = _single(d, awpsgt.AWP_STORAGE_FP32, nsteps=nt)
        a, c = os.path.join(str(tmp_path), "a%d" % k), os.path.join(str(tmp_path), "c%d" % k)
        r.write_sgt_file(a)
        ref.write_sgt_file(c)
        assert open(a, "rb").read() == open(c, "rb").read(), d
        ref.close()
        r.close()


def test_batch_rejects_bad_runs(case_dirs):
    d = case_dirs["g1_fx_layered"][0]
    r1 = _make(d, awpsgt.AWP_STORAGE_FP32)
    r2 = _make(d, awpsgt.AWP_STORAGE_MIXED)
    r3 = _make(d, awpsgt.AWP_STORAGE_FP32, init=False)
    with pytest.raises(awpsgt.AwpError, match="storage"):
        awpsgt.Batch([r1, r2])
    with pytest.raises(awpsgt.AwpError, match="initialised"):
        awpsgt.Batch([r1, r3])
    with pytest.raises(awpsgt.AwpError, match="twice"):
        awpsgt.Batch([r1, r1])
    with awpsgt.Batch([r1]) as b:
        r1.step(3)
        with pytest.raises(awpsgt.AwpError, match="not stepped"):
            b.tune(1)
    for r in (r1, r2, r3):
        r.close()


def test_estimate_bytes_matches_allocation(case_dirs):
    for storage in (awpsgt.AWP_STORAGE_FP32, awpsgt.AWP_STORAGE_MIXED):
        d = case_dirs["g7_random_edges"][0]
        r = _make(d, storage)
        steady, peak = awpsgt.estimate_bytes(r.p, storage, r.npts)
        held = r.device_bytes()
        assert held <= steady <= held + 64 * 1024, (storage, held, steady)
        assert peak > steady
        r.close()


def test_host_batch_mode_writes_each_runs_file(case_dirs, tmp_path):
    """`awp_pmcl3d --batch IN3D_a IN3D_b ...` (one process, one wave, one
    batch) writes for every run the file the single-run host writes."""
    import shutil
    import subprocess
    from conftest import REPO
    exe = os.path.join(REPO, "cybershake-core_amd", "bin", "awp_pmcl3d")
    run = str(tmp_path)
    names = ["g1_fx_layered", "g2_fy_nvar5_nosocalq", "g3_fz_surface"]
    in3ds = []
    for k, name in enumerate(names):
        sub = os.path.join(run, "r%d" % k)
        shutil.copytree(case_dirs[name][0], sub)
        txt = open(os.path.join(sub, "IN3D")).read()
        for f in ("source", "media", "cordfile", "sgt", "CHKP", "CHKJ"):
            txt = txt.replace("'%s'" % f, "'r%d/%s'" % (k, f))
        open(os.path.join(run, "IN3D.%d" % k), "w").write(txt)
        in3ds.append("IN3D.%d" % k)
    env = dict(os.environ, AWP_SGT_RWG="rwg")
    r = subprocess.run([exe, "--batch"] + in3ds, cwd=run, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout
    assert "wave 0: runs 0..2" in r.stdout
    for k, name in enumerate(names):
        ref = _single(case_dirs[name][0], awpsgt.AWP_STORAGE_FP32)
        want = os.path.join(run, "want%d" % k)
        ref.write_sgt_file(want)
        ref.write_sgt_file_rwg(want + ".rwg", ref.p.igreen == 3)
        ref.close()
        assert open(os.path.join(run, "r%d" % k, "sgt"), "rb").read() == open(want, "rb").read(), name
        assert open(os.path.join(run, "rwg.%d" % k), "rb").read() == open(want + ".rwg", "rb").read(), name


def test_two_stage_run_after_fused_batch(case_dirs, monkeypatch):
    """A run created with both wavefield buffers but switched to the
    two-stage step, advanced an odd number of steps by a fused batch (which
    leaves its state in buffer 1), then stepped on its own: the two-stage
    step must continue from buffer 1's state, bit for bit as a two-stage run
    stepped alone."""
    d = case_dirs["g1_fx_layered"][0]
    monkeypatch.setenv("AWP_SCHEME", "fused")
    r = _make(d, awpsgt.AWP_STORAGE_FP32)
    monkeypatch.delenv("AWP_SCHEME", raising=False)
    r.set_scheme(False)
    assert r.scheme()[0] is False
    monkeypatch.setenv("AWP_BATCH_SCHEME", "fused")
    with awpsgt.Batch([r]) as b:
        assert b.fused()
        b.step(7)
    monkeypatch.delenv("AWP_BATCH_SCHEME", raising=False)
    assert r.progress()[0] == 7
    r.step(r.p.nt)
    r.synchronize()
    monkeypatch.setenv("AWP_SCHEME", "split")
    ref = _single(d, awpsgt.AWP_STORAGE_FP32)
    monkeypatch.delenv("AWP_SCHEME", raising=False)
    assert np.array_equal(r.sgt(), ref.sgt())
    for f in ("u1", "v1", "w1", "xx", "zz", "xz", "yz", "r1", "r6"):
        k = 2 if f in ("u1", "w1", "xz") else 0
        assert np.array_equal(r.field(f, k), ref.field(f, k)), f
    ref.close()
    r.close()


@pytest.mark.parametrize("scheme", ["split", "fused"])
@pytest.mark.parametrize("name", ["g4_surface_xz", "g5_surface_yz", "g6_surface_zz"])
def test_batch_surface_source_sampled_at_source_cell(case_dirs, name, scheme, monkeypatch, tmp_path):
    """Surface sources (igreen 4..6) write stresses at k = nz; SGT points on
    the source's column at the surface (and just below) read them.  The batch
    applies the source sample before taking the SGT sample, as the single run
    and the reference do (pmcl3d.f:1381-1383): outputs equal the single run's
    bit for bit, with several runs in the batch."""
    spec, media, pts = cases.tiny_inputs(name)
    sx, sy, _ = spec.src
    extra = np.array([[sx, sy, 1], [sx, sy, 2], [sx + 1, sy, 1], [sx, sy + 1, 1]], dtype=np.int64)
    pts = np.concatenate([extra, pts])
    d = str(tmp_path / name)
    cases.write_case(d, spec, media, pts)
    dirs = [d, case_dirs["g1_fx_layered"][0], case_dirs[name][0]]
    monkeypatch.setenv("AWP_SCHEME", scheme)
    runs = [_make(x, awpsgt.AWP_STORAGE_FP32) for x in dirs]
    with awpsgt.Batch(runs) as b:
        assert b.fused() == (scheme == "fused")
        b.step(runs[0].p.nt)
    for x, r in zip(dirs, runs):
        ref = _single(x, awpsgt.AWP_STORAGE_FP32)
        assert np.array_equal(r.sgt(), ref.sgt()), x
        ref.close()
        r.close()
    # the surface points really see the source
    monkeypatch.setenv("AWP_SCHEME", "split")
    ref = _single(d, awpsgt.AWP_STORAGE_FP32)
    s = ref.sgt()
    assert np.any(s[0, :, :] != 0) and np.all(np.isfinite(s[0]))
    ref.close()


@pytest.mark.parametrize("storage", [awpsgt.AWP_STORAGE_FP32])
def test_batched_runs_vs_reference_files(case_dirs, storage, tmp_path):
    """All seven seeded cases in one batch (g7 has more steps: it finishes
    on its own after the batch): each run's raw SGT file against the
    reference binary's file, per component (rel L2 <= 1e-4) and per
    (point, component) series above the floor (rel L2 <= 1e-4)."""
    from conftest import load_golden, rel_l2
    from test_sgtcheck import SERIES_FLOOR, per_series_check
    names = list(cases.CASES)
    runs = [_make(case_dirs[n][0], storage) for n in names]
    with awpsgt.Batch(runs) as b:
        b.step(10_000)
    worst = 0.0
    for n, r in zip(names, runs):
        if r.progress()[0] < r.p.nt:
            r.step(r.p.nt)
        out = os.path.join(str(tmp_path), n)
        r.write_sgt_file(out)
        mine = np.fromfile(out, dtype=np.float32)
        ref = load_golden(n)["sgt"]
        assert mine.size == ref.size, n
        nt_out, _ = awpsgt.sgt_file_geometry(r.p)
        npad = (-ref.size) % (6 * nt_out)
        m = np.concatenate([mine, np.zeros(npad, np.float32)]).reshape(-1, 6, nt_out)
        g = np.concatenate([ref, np.zeros(npad, np.float32)]).reshape(-1, 6, nt_out)
        for c in range(6):
            assert rel_l2(m[:, c], g[:, c]) <= 1e-4, (n, c)
        cmp_ = per_series_check(ref, mine, nt_out, tmp_path, n)
        assert cmp_["series_checked"] > 0 and cmp_["worst_rel_l2"] <= 1e-4, (n, cmp_)
        worst = max(worst, cmp_["worst_rel_l2"])
        r.close()
    print("batched runs vs reference files: worst per-series rel L2 %.3g (floor %.0e)" % (worst, SERIES_FLOOR))
