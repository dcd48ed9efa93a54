"""The CPU restatement (oracle/awp_oracle.c) pinned against the reference
solver's own outputs (tests/golden, written by the reference Fortran binary)
and against closed forms that a plausible slip would violate."""
import math
import os

import numpy as np
import pytest

from conftest import load_golden, rel_l2
import cases
from oracle import oracle as orc


def _oracle_for(d, dtype=np.float32):
    p = orc.read_in3d(os.path.join(d, "IN3D"))
    media = orc.read_media(os.path.join(d, "media"), p["NX"], p["NY"], p["NZ"], p["NVAR"])
    xyz, stf = orc.read_source(os.path.join(d, "source"), p["NSRC"], p["NST"])
    pts = orc.read_cordfile(os.path.join(d, "cordfile"))
    return p, media, xyz, stf, pts, orc.Oracle(p, media, xyz, stf, pts, dtype=dtype)


@pytest.mark.parametrize("name", list(cases.CASES))
def test_inputs_match_fixture_generation(case_dirs, name):
    d, sha = case_dirs[name]
    assert str(load_golden(name)["inputs_sha256"]) == sha, "input generator drifted from the fixture"


@pytest.mark.parametrize("name", list(cases.CASES))
def test_oracle_vs_reference_binary(case_dirs, name):
    """FP32 restatement vs the reference's raw SGT file: same size (point
    selection, dropped points, flush geometry) and per-component rel L2 at
    rounding level (the two differ only in libm's exp/log ulps)."""
    d, _ = case_dirs[name]
    p, media, xyz, stf, pts, o = _oracle_for(d)
    o.run(p["NT"])
    s, _ = o.sgt()
    kidx, _ = o.points()
    mine = orc.sgt_file_from_samples(s, p["NT"], p["NTISKP_SGT"], p["WRITE_STEP"], kidx)
    ref = load_golden(name)["sgt"]
    mine = np.frombuffer(mine, dtype=np.float32)
    assert mine.size == ref.size
    nt_out, _ = orc.sgt_file_layout(p["NT"], p["NTISKP_SGT"], p["WRITE_STEP"])
    npad = (-ref.size) % (6 * nt_out)
    r = np.concatenate([ref, np.zeros(npad, np.float32)]).reshape(-1, 6, nt_out)
    m = np.concatenate([mine, np.zeros(npad, np.float32)]).reshape(-1, 6, nt_out)
    for c in range(6):
        assert rel_l2(m[:, c], r[:, c]) < 1e-6, (name, c)


@pytest.mark.slow
def test_oracle_vs_reference_c1(c1_dir):
    """200x200x100, 1500 steps, 2000 points against the reference run's
    subset and per-component norms (about 2 minutes on 8 cores)."""
    d, sha = c1_dir
    g = load_golden("c1_subset")
    assert str(g["inputs_sha256"]) == sha
    p, media, xyz, stf, pts, o = _oracle_for(d)
    o.run(p["NT"], 8)
    s, n = o.sgt()
    assert n == 150
    for c in range(6):
        assert rel_l2(s[g["idx"], c], g["sgt"][:, c]) < 1e-5
        assert abs(np.linalg.norm(s[:, c].astype(np.float64)) / g["comp_l2"][c] - 1) < 1e-5


def test_media_transform_closed_form(case_dirs):
    """d1, mu, lam, qp, qs follow structure.f:449-524, evaluated here in FP64
    from the formulas (dispersion correction, compliances, Q weights)."""
    d, _ = case_dirs["g1_fx_layered"]
    p, media, xyz, stf, pts, o = _oracle_for(d)
    vp = media[..., 0].astype(np.float64)[::-1]  # k = 1 is the bottom plane
    vs = media[..., 1].astype(np.float64)[::-1]
    rho = media[..., 2].astype(np.float64)[::-1]
    qs = 0.05 * vs
    qp = 2 * qs
    w0, w1, w2 = 2 * math.pi * p["FP"], 2 * math.pi * p["FL"], 2 * math.pi * p["FH"]
    vs2 = vs * (1 + math.log(w2 / w0) / (math.pi * qs))
    vp2 = vp * (1 + math.log(w2 / w0) / (math.pi * qp))
    vs2 = np.where(vp2 / vs2 < 1.45, vp2 / 1.45, vs2)
    tmax, tmin = 1 / w1, 1 / w2
    t1 = 2 / math.pi * math.log(tmax / tmin)
    t2 = 2 / math.pi * math.log(w0 * tmin)
    exp = {"d1": rho, "mu": 1 / (rho * vs2 ** 2), "lam": 1 / (rho * (vp2 ** 2 - 2 * vs2 ** 2)),
           "qp": t1 / qp / (1 - t2 / qp), "qs": t1 / qs / (1 - t2 / qs)}
    for k, v in exp.items():
        got = o.field(k).astype(np.float64)
        assert np.max(np.abs(got / v - 1)) < 1e-5, k


def test_sponge_and_relaxation_closed_form():
    """Cerjan factors exp(-(alpha*(ND-n+1))^2) with alpha = sqrt(-ln ARBC)/ND
    (cerjan.f:25-41) on both x faces, both y faces and the bottom; the eight
    relaxation times are geometrically spaced between tau_min and tau_max with
    the 2x2x2 pattern of relax.f:15-34."""
    nd, arbc, n = 20, 0.92, 200
    a = math.sqrt(-math.log(arbc)) / nd
    f = [1.0] * (n + 1)
    for i in range(1, n + 1):
        if i <= nd:
            f[i] = math.exp(-(a * (nd - i + 1)) ** 2)
        elif i >= n - nd + 1:
            f[i] = math.exp(-(a * (nd - (n - i))) ** 2)
    # symmetric, minimum ARBC^(1) at the faces... exp(-(a*nd)^2) = ARBC
    assert abs(f[1] - arbc) < 1e-12 and abs(f[n] - arbc) < 1e-12
    assert f[nd + 1] == 1.0 and f[n - nd] == 1.0
    assert all(f[i] < f[i + 1] for i in range(1, nd))
    # relaxation pattern: the 8 times span [tau_min, tau_max] geometrically
    fl, fh = 0.01, 25.0
    tmax, tmin = 1 / (2 * math.pi * fl), 1 / (2 * math.pi * fh)
    pattern = {(0, 0, 0): 1, (1, 0, 0): 6, (0, 1, 0): 7, (1, 1, 0): 4,
               (0, 0, 1): 8, (1, 0, 1): 3, (0, 1, 1): 2, (1, 1, 1): 5}
    taus = {k: math.exp(0.5 * (math.log(tmax * tmin) + math.log(tmax / tmin) * (2 * (t - 0.5) / 8 - 1)))
            for k, t in pattern.items()}
    vals = sorted(taus.values())
    ratios = [vals[q + 1] / vals[q] for q in range(7)]
    assert max(ratios) / min(ratios) - 1 < 1e-12
    assert tmin < vals[0] < vals[-1] < tmax


def linearity_check(s1, s2):
    """Doubling the source doubles the SGT.  Not bit-for-bit: x2 commutes with
    rounding only for normal floats, and the leading edge of the wavefront
    passes through the subnormal range, where the rounding grid is absolute;
    the relative error seeded there travels with the wave.  The whole-run
    rel L2 stays at ~1e-9."""
    a, b = 2 * s1.astype(np.float64), s2.astype(np.float64)
    assert np.count_nonzero(a) > 100
    assert rel_l2(b, a) < 1e-6


def test_linearity(case_dirs):
    """Every update is linear in the wavefield with media-only coefficients."""
    d, _ = case_dirs["g1_fx_layered"]
    p, media, xyz, stf, pts, o1 = _oracle_for(d)
    o2 = orc.Oracle(p, media, xyz, 2 * stf, pts)
    o1.run(60)
    o2.run(60)
    s1, _ = o1.sgt()
    s2, _ = o2.sgt()
    linearity_check(s1[:, :, :6], s2[:, :, :6])


def test_zero_source_zero_field(case_dirs):
    d, _ = case_dirs["g7_random_edges"]
    p, media, xyz, stf, pts, _o = _oracle_for(d)
    o = orc.Oracle(p, media, xyz, 0 * stf, pts)
    o.run(30)
    for f in ("u1", "w1", "xx", "yz", "r3"):
        assert not np.any(o.field(f))


def test_fp64_companion_close_to_fp32(case_dirs):
    d, _ = case_dirs["g2_fy_nvar5_nosocalq"]
    p, media, xyz, stf, pts, o32 = _oracle_for(d)
    o64 = orc.Oracle(p, media, xyz, stf, pts, dtype=np.float64)
    o32.run(p["NT"])
    o64.run(p["NT"])
    a, _ = o32.sgt()
    b, _ = o64.sgt()
    for c in range(6):
        assert rel_l2(a[:, c], b[:, c]) < 1e-4


def test_free_surface_images(case_dirs):
    """After a step: xz = yz = 0 on the surface and the stress images are odd
    reflections (operator.f:575-612); u(nz+2) stays zero."""
    d, _ = case_dirs["g3_fz_surface"]
    p, media, xyz, stf, pts, o = _oracle_for(d)
    o.run(25)
    nz = p["NZ"]
    zz, xz, yz = o.field("zz", 2), o.field("xz", 2), o.field("yz", 2)
    assert not np.any(xz[nz - 1]) and not np.any(yz[nz - 1])
    assert np.array_equal(zz[nz], -zz[nz - 1]) and np.array_equal(zz[nz + 1], -zz[nz - 2])
    assert np.array_equal(xz[nz], -xz[nz - 2]) and np.array_equal(xz[nz + 1], -xz[nz - 3])
    assert not np.any(o.field("u1", 2)[nz + 1])


@pytest.mark.parametrize("name", ["g1_fx_layered", "g7_random_edges"])
def test_mixed_media_storage_closed_form(case_dirs, name):
    """The mixed restatement stores d1, mu, lam as FP16 deviations h from a
    per-level profile b: b = the midrange of the level's nonzero interior
    values (double, rounded to float), h = FP16 of (float)(x / b - 1) (round
    to nearest even), x' = fma(b, h, b).  Derived here in numpy from the FP32
    restatement's values: bit for bit; exact on a layered medium; within
    2^-12 b of x everywhere."""
    p, media, xyz, stf, pts, o32 = _oracle_for(case_dirs[name][0])
    om = orc.Oracle(p, media, xyz, stf, pts, mixed=True)
    for f in ("d1", "mu", "lam"):
        x = o32.field(f)                       # [k][j][i], k = 1 at the bottom
        got = om.field(f)
        nz_ = x != 0
        lo = np.where(nz_, x, np.inf).min(axis=(1, 2)).astype(np.float64)
        hi = np.where(nz_, x, -np.inf).max(axis=(1, 2)).astype(np.float64)
        b = ((lo + hi) * 0.5).astype(np.float32)[:, None, None]
        h = (x.astype(np.float64) / b.astype(np.float64) - 1.0).astype(np.float32).astype(np.float16)
        want = (b.astype(np.float64) * h.astype(np.float64) + b.astype(np.float64)).astype(np.float32)
        assert np.array_equal(got, want), f
        err = np.abs(got.astype(np.float64) - x)
        assert np.all(err <= b.astype(np.float64) * 2.0 ** -12 * (1 + 2.0 ** -10) + np.abs(x) * 2.0 ** -24), f
        if name == "g1_fx_layered":
            assert np.array_equal(got, x), f
        else:
            assert not np.array_equal(got, x), f
    o32.close()
    om.close()
