"""Mixed-storage mode and the FP16 SGT file, through the C ABI.

Mixed storage keeps the wavefield, density and compliances in FP32 and
stores the memory variables r1..r6 as FP16 under a power-of-two exponent per
64-point x row segment, and the attenuation weights qp/qs
as FP16 (value * 2^e); arithmetic stays FP32.  The reference has no such
mode, so the checks are:
  * bit-for-bit equality with the oracle's restatement of the mixed mode
    (same rounding points, same operation order);
  * SGT error against the FP64 oracle run <= 1e-3 relative L2 per
    component (the bound the mixed mode is built to), with the FP32 device
    run's own error against FP64 printed beside it;
  * the FP16 SGT file decodes to the FP32 samples rounded to FP16 after an
    exact power-of-two scaling per point and component, so every decoded
    value is within 2^-11 of its series' peak magnitude.
"""
import os
import struct

import numpy as np
import pytest

from conftest import rel_l2
import cases
import awpsgt
from oracle import oracle as orc
from test_gpu_parity import _inputs, _oracle

pytestmark = pytest.mark.gpu

MIX_TOL = 1e-3      # mixed storage vs FP64, rel L2 per SGT component
FP32_TOL = 1e-4     # FP32 vs FP64


def _run(p, media, xyz, stf, pts, storage, nsteps=None):
    r = awpsgt.Run(p, storage=storage)
    r.set_media(media)
    r.set_source(xyz, stf)
    r.set_sgt_points(pts)
    r.init()
    r.step(p.nt if nsteps is None else nsteps)
    r.synchronize()
    return r


def _mixed_oracle(d, media, xyz, stf, pts, dtype=np.float32, mixed=True):
    q = orc.read_in3d(os.path.join(d, "IN3D"))
    return orc.Oracle(q, media, xyz, stf, pts, dtype=dtype, mixed=mixed)


@pytest.mark.parametrize("name", ["g7_random_edges", "g1_fx_layered"])
def test_mixed_media_bitwise(case_dirs, name):
    """Stored media after set-up equal the mixed restatement's bit for bit:
    d1/mu/lam as FP16 deviations from their level profiles (exact on a
    layered medium, within FP16's rounding of x / b - 1 on a random one), and
    qp/qs as FP16 values times their scale."""
    d, _ = case_dirs[name]
    p, media, xyz, stf, pts = _inputs(d)
    with _run(p, media, xyz, stf, pts, awpsgt.AWP_STORAGE_MIXED, nsteps=0) as r:
        o = _mixed_oracle(d, media, xyz, stf, pts)
        o32 = _oracle(d, p, media, xyz, stf, pts)
        for f in ("d1", "mu", "lam", "qp", "qs"):
            assert np.array_equal(r.field(f), o.field(f)), f
        for f in ("d1", "mu", "lam"):
            a, b = r.field(f).astype(np.float64), o32.field(f).astype(np.float64)
            if name == "g1_fx_layered":
                assert np.array_equal(a, b), f           # one value per level: h = 0
            else:
                assert not np.array_equal(a, b), f       # really FP16
                # in the interior |h| < 1 (b is the level's midrange), so the
                # FP16 rounding of h moves x by at most 2^-12 b <= 2^-12 of the
                # level's largest value, and the fma rounds once more
                lvmax = np.abs(b).max(axis=(1, 2), keepdims=True)
                assert np.all(np.abs(a - b) <= lvmax * 2.0 ** -12 + np.abs(b) * 2.0 ** -23), f
        for f in ("qp", "qs"):
            a, b = r.field(f), o32.field(f)
            assert not np.array_equal(a, b)          # really FP16
            assert np.max(np.abs(a - b)) <= np.max(np.abs(b)) * 2.0 ** -11


@pytest.mark.parametrize("scheme", ["split", "fused", "split_o32", "split_o64", "split_ty16",
                                    "split_xw2", "split_xw2_ty4", "split_xw2_o64", "split_o32v",
                                    "split_o64v", "fused_ty11", "fused_lv2"])
@pytest.mark.parametrize("name", list(cases.CASES))
def test_mixed_fields_bitwise_after_steps(case_dirs, name, scheme, monkeypatch):
    d, _ = case_dirs[name]
    p, media, xyz, stf, pts = _inputs(d)
    o = _mixed_oracle(d, media, xyz, stf, pts)
    monkeypatch.setenv("AWP_SCHEME", "fused" if scheme.startswith("fused") else "split")
    if scheme in ("fused_ty11", "fused_lv2"):
        # 64 x 11 tiles: 16-wave blocks
        monkeypatch.setenv("AWP_TILE_Y", "11")
    if scheme in ("fused_ty11", "fused_lv2"):
        # one level per barrier, or two (the 64 x 11 default)
        monkeypatch.setenv("AWP_PF", "0" if scheme == "fused_ty11" else "3")
    if scheme in ("split_o32", "split_o64"):
        # both byte-offset widths of the two-stage kernels, whichever the
        # storage mode uses by default
        monkeypatch.setenv("AWP_O32", "1" if scheme == "split_o32" else "0")
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
    with _run(p, media, xyz, stf, pts, awpsgt.AWP_STORAGE_MIXED, nsteps=0) as r:
        for target in (1, 9, 40):
            r.step(target - r.progress()[0])
            o.run(target - o.info()["step"])
            for f in awpsgt.FIELDS[:15]:
                k = 2 if f in ("u1", "v1", "w1", "zz", "xz", "yz") else 0
                a, b = r.field(f, k), o.field(f, k)
                assert np.array_equal(a, b), (name, target, f, float(np.abs(a - b).max()))
        for f in ("r1", "r4", "r6"):
            a = r.field(f)
            # FP16 values times powers of two: 11 significant bits at most
            assert np.all((a.view(np.uint32) & 0x1FFF) == 0), f


@pytest.mark.parametrize("name", list(cases.CASES))
def test_mixed_sgt_vs_fp64(case_dirs, name, tmp_path):
    d, _ = case_dirs[name]
    p, media, xyz, stf, pts = _inputs(d)
    with _run(p, media, xyz, stf, pts, awpsgt.AWP_STORAGE_MIXED) as r:
        out = os.path.join(str(tmp_path), "sgt_mixed")
        r.write_sgt_file(out)
        mine = open(out, "rb").read()
        sm = r.sgt()
    om = _mixed_oracle(d, media, xyz, stf, pts)
    om.run(p.nt)
    s, _ = om.sgt()
    kidx, _ = om.points()
    assert mine == orc.sgt_file_from_samples(s, p.nt, p.ntiskp_sgt, p.write_step, kidx)
    with _run(p, media, xyz, stf, pts, awpsgt.AWP_STORAGE_FP32) as r32:
        s32 = r32.sgt()
    o64 = _mixed_oracle(d, media, xyz, stf, pts, dtype=np.float64, mixed=False)
    o64.run(p.nt)
    s64, n = o64.sgt()
    kept = kidx
    em, e32 = [], []
    for c in range(6):
        ref = s64[kept, c, :n]
        em.append(rel_l2(sm[kept, c, :n], ref))
        e32.append(rel_l2(s32[kept, c, :n], ref))
    print("%s mixed vs FP64: %s; FP32 vs FP64: %s" % (
        name, " ".join("%.2e" % e for e in em), " ".join("%.2e" % e for e in e32)))
    assert max(em) <= MIX_TOL, em
    assert max(e32) <= FP32_TOL, e32


def _expected_fp16(samples):
    """FP16 encoding of float32 series [..., nsamp]: max|x| = f * 2^e',
    f in [0.5, 1), e = max(e', -134); scale 2^(e-15); data = fp16(x *
    2^(15-e)), round to nearest even (numpy's float32 -> float16 cast);
    all-zero series: scale 0, data 0."""
    mx = np.max(np.abs(samples), axis=-1)
    _, ex = np.frexp(mx.astype(np.float32))
    ex = np.maximum(ex, -134)
    scale = np.where(mx > 0, np.ldexp(np.float32(1), ex - 15), 0).astype(np.float32)
    scaled = np.ldexp(samples.astype(np.float32), (15 - ex)[..., None]).astype(np.float32)
    data = np.where(mx[..., None] > 0, scaled, 0).astype(np.float16)
    return scale, data


@pytest.mark.parametrize("name", ["g1_fx_layered", "g7_random_edges"])
def test_sgt_fp16_file(case_dirs, name, tmp_path):
    d, _ = case_dirs[name]
    p, media, xyz, stf, pts = _inputs(d)
    with _run(p, media, xyz, stf, pts, awpsgt.AWP_STORAGE_FP32) as r:
        path = os.path.join(str(tmp_path), "sgt16")
        r.write_sgt_file_fp16(path)
        s = r.sgt()
        keep, _ = awpsgt.sgt_select(p, pts)
    raw = open(path, "rb").read()
    assert raw[:8] == b"AWPSGT16"
    version, npts, ncomp, nsamp, ntiskp = struct.unpack("<5I", raw[8:28])
    (dt,) = struct.unpack("<f", raw[28:32])
    (nto,) = struct.unpack("<I", raw[32:36])
    # the samples the raw file holds: full WRITE_STEP buffers only
    nt_out, nflush = awpsgt.sgt_file_geometry(p)
    assert nto == nt_out
    assert (version, npts, ncomp, nsamp, ntiskp) == (1, len(pts), 6, nflush * p.write_step, p.ntiskp_sgt)
    assert dt == np.float32(p.dt)
    assert len(raw) == 64 + npts * 6 * 4 + npts * 6 * nsamp * 2
    scale = np.frombuffer(raw, dtype=np.float32, count=npts * 6, offset=64).reshape(npts, 6)
    data = np.frombuffer(raw, dtype=np.float16, offset=64 + npts * 6 * 4).reshape(npts, 6, nsamp)
    keep = np.asarray(keep, dtype=bool)
    x = s[:, :, :nsamp]
    es, ed = _expected_fp16(x[keep])
    assert np.array_equal(scale[keep], es)
    # equal as FP16 values (a -0.0 sample may come out as +0.0)
    got, want = data[keep], ed
    bad = np.argwhere(got != want)
    for b in bad[:8]:
        b = tuple(b)
        xs = x[keep][b]
        print("mismatch at %s: x=%r (bits %08x) x/scale=%r want %04x got %04x"
              % (b, float(xs), np.float32(xs).view(np.uint32), float(xs / es[b[:2]]), want.view(np.uint16)[b], got.view(np.uint16)[b]))
    assert len(bad) == 0, "%d of %d FP16 values differ" % (len(bad), got.size)
    # points the rule drops: zero scale and data
    assert np.all(scale[~keep] == 0) and np.all(data[~keep].view(np.uint16) == 0)
    dec = awpsgt.read_sgt16(path)
    assert dec.shape == (npts, 6, nsamp)
    peak = np.max(np.abs(x[keep]), axis=-1, keepdims=True).astype(np.float64)
    err = np.abs(dec[keep].astype(np.float64) - x[keep])
    assert np.all(err <= peak * 2.0 ** -11 + 2.0 ** -149)
    assert np.array_equal(dec[keep], ed.astype(np.float32) * es[..., None])
    if name == "g7_random_edges":
        assert not keep.all()    # the case holds points the rule drops
    # DirectSynth's view: RWG records decoded from the FP16 file match the
    # reformatted FP32 raw file to FP16 rounding
    from oracle.sgthead import reformat_awp
    rawf = os.path.join(str(tmp_path), "raw")
    with _run(p, media, xyz, stf, pts, awpsgt.AWP_STORAGE_FP32) as r2:
        r2.write_sgt_file(rawf)
    rwg = np.frombuffer(reformat_awp(open(rawf, "rb").read(), nt_out), "<f4").reshape(-1, 6, nt_out)
    dec = awpsgt.read_sgt16_points(path, 0, rwg.shape[0], nt_out, 1)
    pk = np.max(np.abs(rwg), axis=-1, keepdims=True).astype(np.float64)
    assert np.all(np.abs(dec.astype(np.float64) - rwg) <= pk * 2.0 ** -11 + 2.0 ** -149)
