"""Parity of the HIP path (through the C ABI) with the CPU restatement and
the reference solver's own outputs.  FP32 device results are compared
bit-for-bit with the FP32 restatement (same IEEE operation sequence, no FMA
contraction) and within rel L2 <= 1e-4 per SGT component with the reference
binary's files (the task's FP32 tolerance)."""
import hashlib
import os
import subprocess

import numpy as np
import pytest

from conftest import REPO, load_golden, rel_l2
import cases
import awpsgt
import awpsgt.inputs
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _inputs(d):
    p = awpsgt.read_in3d(os.path.join(d, "IN3D"))
    media = awpsgt.read_media(os.path.join(d, "media"), p)
    xyz, stf = awpsgt.read_source(os.path.join(d, "source"), p.nsrc, p.nst)
    pts = awpsgt.read_cordfile(os.path.join(d, "cordfile"))
    return p, media, xyz, stf, pts


def _gpu_run(p, media, xyz, stf, pts, nsteps=None):
    r = awpsgt.Run(p)
    r.set_media(media)
    r.set_source(xyz, stf)
    r.set_sgt_points(pts)
    r.init()
    r.step(p.nt if nsteps is None else nsteps)
    r.synchronize()
    return r


def _oracle(d, p, media, xyz, stf, pts):
    q = orc.read_in3d(os.path.join(d, "IN3D"))
    return orc.Oracle(q, media, xyz, stf, pts)


def test_media_setup_bitwise(case_dirs):
    d, _ = case_dirs["g7_random_edges"]
    p, media, xyz, stf, pts = _inputs(d)
    with _gpu_run(p, media, xyz, stf, pts, nsteps=0) as r:
        o = _oracle(d, p, media, xyz, stf, pts)
        for f in ("d1", "mu", "lam", "qp", "qs"):
            assert np.array_equal(r.field(f), o.field(f)), f


@pytest.mark.parametrize("scheme", ["split", "fused", "split_o32", "split_o64", "split_ty16",
                                    "split_xw2", "split_xw2_ty4", "split_xw2_o64", "split_o32v",
                                    "split_o64v", "fused_ty11", "fused_lv2", "fused_static", "fused_claim_r5",
                                    "split_claim", "split_ty4_kc8"])
@pytest.mark.parametrize("name", list(cases.CASES))
def test_fields_bitwise_after_steps(case_dirs, name, scheme, monkeypatch):
    """Both time-step schemes (AWP_SCHEME=split two-stage, AWP_SCHEME=fused),
    both offset forms of the two-stage stress and velocity kernels (32-bit
    byte offsets, and the 64-bit element indices that grids with arrays of
    4 GiB or more take), their tile shapes (64 x 8, 64 x 16, 128 x 8,
    128 x 4, 64 x 4 in 8-level z chunks) and the static and claimed tile
    assignments (the fused ranges rotated) against the restatement, bit for
    bit."""
    d, _ = case_dirs[name]
    p, media, xyz, stf, pts = _inputs(d)
    o = _oracle(d, p, media, xyz, stf, pts)
    monkeypatch.setenv("AWP_SCHEME", "fused" if scheme.startswith("fused") else "split")
    if scheme in ("fused_ty11", "fused_lv2"):
        # 64 x 11 tiles: 16-wave blocks
        monkeypatch.setenv("AWP_TILE_Y", "11")
    if scheme in ("fused_ty11", "fused_lv2"):
        # one level per barrier, or two (the 64 x 11 default)
        monkeypatch.setenv("AWP_PF", "0" if scheme == "fused_ty11" else "3")
    if scheme in ("fused_static", "fused_claim_r5", "split_claim"):
        # one block per tile, or persistent blocks claiming tiles per XCD
        # (forced: these grids have too few tiles for the default to claim),
        # the fused ranges rotated; each claimed launch leaves the counters
        # zero for the next one
        monkeypatch.setenv("AWP_CLAIM", "0" if scheme == "fused_static" else "2")
        monkeypatch.setenv("AWP_XROT", "5" if scheme == "fused_claim_r5" else "0")
    if scheme in ("split_o32", "split_o64"):
        # both byte-offset widths of the two-stage kernels, whichever the
        # storage mode uses by default
        monkeypatch.setenv("AWP_O32", "1" if scheme == "split_o32" else "0")
    if scheme == "split_ty4_kc8":
        # 64 x 4 tiles in z chunks of 8 levels (a small-grid tuner candidate)
        monkeypatch.setenv("AWP_TILE_Y", "4")
        monkeypatch.setenv("AWP_KC", "8")
    if scheme == "split_ty16":
        # 64 x 16 tiles: the 1024-thread variants of the two-stage kernels
        monkeypatch.setenv("AWP_TILE_Y", "16")
    if scheme in ("split_o32v", "split_o64v"):
        # both forms of the velocity kernel
        monkeypatch.setenv("AWP_O32V", "1" if scheme == "split_o32v" else "0")
    if scheme.startswith("split_xw2"):
        # 128-point-wide tiles (two waves per tile row): 128 x 8 tiles in
        # 1024-thread blocks, 128 x 4 tiles in both stress forms
        monkeypatch.setenv("AWP_TILE_XW", "2")
        if scheme != "split_xw2":
            monkeypatch.setenv("AWP_TILE_Y", "4")
        if scheme == "split_xw2_o64":
            monkeypatch.setenv("AWP_O32", "0")
    with _gpu_run(p, media, xyz, stf, pts, nsteps=0) as r:
        for target in (1, 7, 45):
            r.step(target - r.progress()[0])
            o.run(target - o.info()["step"])
            for f in awpsgt.FIELDS[:15]:
                a, b = r.field(f, 2 if f in ("u1", "v1", "w1", "zz", "xz", "yz") else 0), \
                    o.field(f, 2 if f in ("u1", "v1", "w1", "zz", "xz", "yz") else 0)
                assert np.array_equal(a, b), (name, target, f, float(np.abs(a - b).max()))


@pytest.mark.parametrize("name", list(cases.CASES))
def test_sgt_vs_reference_binary(case_dirs, name, tmp_path):
    d, _ = case_dirs[name]
    p, media, xyz, stf, pts = _inputs(d)
    with _gpu_run(p, media, xyz, stf, pts) as r:
        out = os.path.join(str(tmp_path), "sgt")
        r.write_sgt_file(out)
        mine = np.fromfile(out, dtype=np.float32)
        ref = load_golden(name)["sgt"]
        assert mine.size == ref.size
        nt_out, _ = awpsgt.sgt_file_geometry(p)
        npad = (-ref.size) % (6 * nt_out)
        m = np.concatenate([mine, np.zeros(npad, np.float32)]).reshape(-1, 6, nt_out)
        g = np.concatenate([ref, np.zeros(npad, np.float32)]).reshape(-1, 6, nt_out)
        for c in range(6):
            assert rel_l2(m[:, c], g[:, c]) <= TOL, (name, c)
        # every (point, component) series above the floor (the level where
        # the reference's own rank decompositions agree with each other)
        from test_sgtcheck import per_series_check
        cmp_ = per_series_check(ref, mine, nt_out, tmp_path, name)
        assert cmp_["series_checked"] > 0 and cmp_["worst_rel_l2"] <= TOL, cmp_
        # bit-exact against the FP32 restatement's file
        o = _oracle(d, p, media, xyz, stf, pts)
        o.run(p.nt)
        s, _ = o.sgt()
        kidx, _ = o.points()
        assert mine.tobytes() == orc.sgt_file_from_samples(s, p.nt, p.ntiskp_sgt, p.write_step, kidx)


def test_c1_vs_reference_binary(c1_dir, tmp_path):
    """200x200x100, 1500 steps, 2000 points: every 10th point against the
    reference run, per-component norms over all points, and the SHA-256 of
    the whole file (equal if the run is bit-identical to the reference)."""
    d, sha = c1_dir
    g = load_golden("c1_subset")
    assert str(g["inputs_sha256"]) == sha
    p, media, xyz, stf, pts = _inputs(d)
    with _gpu_run(p, media, xyz, stf, pts) as r:
        out = os.path.join(str(tmp_path), "sgt")
        r.write_sgt_file(out)
        raw = open(out, "rb").read()
        assert len(raw) == int(g["nbytes"])
        a = np.frombuffer(raw, dtype=np.float32).reshape(len(pts), 6, 150)
        worst = 0.0
        for c in range(6):
            e = rel_l2(a[g["idx"], c], g["sgt"][:, c])
            worst = max(worst, e)
            assert e <= TOL, c
            assert abs(np.linalg.norm(a[:, c].astype(np.float64)) / g["comp_l2"][c] - 1) <= TOL
        # per (point, component) series of the subset, above the floor
        from test_sgtcheck import SERIES_FLOOR
        ref = g["sgt"].astype(np.float64)
        mine = a[g["idx"]].astype(np.float64)
        nr = np.linalg.norm(ref, axis=2)
        nd = np.linalg.norm(mine - ref, axis=2)
        keep = nr >= SERIES_FLOOR * nr.max()
        per_series = float((nd[keep] / nr[keep]).max())
        assert keep.sum() > 0.9 * keep.size and per_series <= TOL, per_series
        print("C1 worst per-component rel L2 vs reference: %.3g; worst per-series %.3g; file sha equal: %s"
              % (worst, per_series, hashlib.sha256(raw).hexdigest() == str(g["file_sha256"])))


def test_large_shape_properties():
    """At a large shape (480x400x160): doubling the source doubles the SGT
    (linearity, to rounding), two identical runs agree bit-for-bit
    (no races), and after 8 steps every field equals the CPU restatement
    bit-for-bit."""
    spec = cases.c1_spec()
    spec.nx, spec.ny, spec.nz, spec.nt = 480, 400, 160, 40
    spec.ntiskp_sgt, spec.write_step = 10, 4
    spec.src = (240, 200, 20)
    spec.t0 = 0.05
    media = awpsgt.inputs.layered_media(spec.nx, spec.ny, spec.nz, cases.LAYERS_C1)
    pts = awpsgt.inputs.random_cordfile_points(5000, 7, (1, 480), (1, 400), (1, 160))
    import tempfile
    d = tempfile.mkdtemp()
    cases.write_case(d, spec, media, pts)
    p, media, xyz, stf, pts = _inputs(d)
    with _gpu_run(p, media, xyz, stf, pts) as r1, _gpu_run(p, media, xyz, 2 * stf, pts) as r2:
        s1, s2 = r1.sgt(), r2.sgt()
        assert np.any(s1 != 0) and np.any(s1[:, :, :4] != -1)
        from test_oracle import linearity_check
        linearity_check(s1, s2)
    with _gpu_run(p, media, xyz, stf, pts) as r3:
        assert np.array_equal(r3.sgt(), s1)
    with _gpu_run(p, media, xyz, stf, pts, nsteps=8) as r4:
        o = _oracle(d, p, media, xyz, stf, pts)
        o.run(8, 16)
        for f in ("u1", "w1", "xx", "xz", "yz", "r1", "r5"):
            k = 2 if f in ("u1", "w1", "xz", "yz") else 0
            assert np.array_equal(r4.field(f, k), o.field(f, k)), f


@pytest.mark.parametrize("claim", ["2", "0"])
@pytest.mark.parametrize("storage", [awpsgt.AWP_STORAGE_FP32, awpsgt.AWP_STORAGE_MIXED])
@pytest.mark.parametrize("scheme", ["split", "fused"])
def test_ragged_shape_equals_restatement(scheme, storage, claim, monkeypatch, tmp_path):
    """A ragged shape (200x130x60: rows of three full 64-point segments and
    one of 8 points, partial tiles in y for both tile heights, a random
    medium) with every launch claiming its tiles per XCD (AWP_CLAIM=2; the
    default only claims on grids of many tiles) or one block per tile, in
    both schemes and storage modes: after 10 steps every field equals the
    restatement's (FP32, or the mixed-storage one) bit for bit."""
    spec = cases.c1_spec()
    spec.nx, spec.ny, spec.nz, spec.nt = 200, 130, 60, 20
    spec.ntiskp_sgt, spec.write_step = 5, 2
    spec.src = (101, 66, 12)
    spec.t0 = 0.05
    media = awpsgt.inputs.random_media(spec.nx, spec.ny, spec.nz, seed=11, vs_top=800.0, vs_bot=3500.0,
                                       pert=0.1)
    pts = awpsgt.inputs.random_cordfile_points(800, 3, (1, 200), (1, 130), (1, 60))
    d = str(tmp_path)
    cases.write_case(d, spec, media, pts)
    p, media, xyz, stf, pts = _inputs(d)
    monkeypatch.setenv("AWP_SCHEME", scheme)
    monkeypatch.setenv("AWP_CLAIM", claim)
    q = orc.read_in3d(os.path.join(d, "IN3D"))
    o = orc.Oracle(q, media, xyz, stf, pts, mixed=storage == awpsgt.AWP_STORAGE_MIXED)
    o.run(10, 20)
    r = awpsgt.Run(p, storage=storage)
    try:
        r.set_media(media)
        r.set_source(xyz, stf)
        r.set_sgt_points(pts)
        r.init()
        assert r.scheme()[0] == (scheme == "fused")
        r.step(10)
        r.synchronize()
        for f in ("u1", "v1", "w1", "xx", "yy", "zz", "xy", "xz", "yz", "r1", "r4", "r6"):
            k = 2 if f in ("u1", "v1", "w1", "zz", "xz", "yz") else 0
            assert np.array_equal(r.field(f, k), o.field(f, k)), f
        assert np.array_equal(r.sgt()[:, :, :2], o.sgt()[0][:, :, :2])
    finally:
        r.close()


def test_host_executable_matches_reference(case_dirs, tmp_path):
    """The C host awp_pmcl3d run on an IN3D directory writes the same file
    as the reference binary (rel L2 per component) and as the library."""
    name = "g7_random_edges"
    d, _ = case_dirs[name]
    import shutil
    run = os.path.join(str(tmp_path), "run")
    shutil.copytree(d, run)
    exe = os.path.join(REPO, "cybershake-core_amd", "bin", "awp_pmcl3d")
    r = subprocess.run([exe, "IN3D"], cwd=run, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    assert "computing time per time step" in r.stdout
    mine = np.fromfile(os.path.join(run, "sgt"), dtype=np.float32)
    ref = load_golden(name)["sgt"]
    assert mine.size == ref.size
    assert rel_l2(mine, ref) <= TOL
    assert os.path.exists(os.path.join(run, "CHKP"))
