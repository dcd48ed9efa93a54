"""Host side of the drop-in surface (no GPU): the C ABI library loads and
exports every declared symbol; the C parsers of IN3D, source and cordfile
agree with the oracle's readers; the SGT point rule and file geometry agree
with the reference (as restated by the oracle and recorded in the fixtures);
unsupported settings are rejected."""
import os
import re

import numpy as np
import pytest

from conftest import REPO, load_golden
import cases
import awpsgt
from awpsgt.inputs import write_in3d
from oracle import oracle as orc


def test_header_symbols_exported():
    hdr = open(os.path.join(REPO, "include", "awpsgt.h")).read()
    declared = set(re.findall(r"\b(awp_[a-z0-9_]+)\s*\(", hdr))
    assert declared == set(awpsgt.EXPORTS)
    L = awpsgt.lib()
    for name in declared:
        assert hasattr(L, name), name


@pytest.mark.parametrize("name", list(cases.CASES))
def test_in3d_source_cordfile_parsers(case_dirs, name):
    d, _ = case_dirs[name]
    p = awpsgt.read_in3d(os.path.join(d, "IN3D"))
    awpsgt.validate_in3d(p)
    q = orc.read_in3d(os.path.join(d, "IN3D"))
    for key in ("IGREEN", "NX", "NY", "NZ", "NPX", "NPY", "NPZ", "ND", "NST", "NSRC",
                "NTISKP_SGT", "WRITE_STEP", "NVAR", "SOCALQ", "NT"):
        assert getattr(p, key.lower()) == q[key], key
    for key in ("DH", "DT", "ARBC", "FL", "FH", "FP", "TMAX"):
        assert np.float32(getattr(p, key.lower())) == np.float32(q[key]), key
    assert p.s("insgt") == "cordfile" and p.s("sgtgro") == "sgt"
    xyz, stf = awpsgt.read_source(os.path.join(d, "source"), p.nsrc, p.nst)
    xyz2, stf2 = orc.read_source(os.path.join(d, "source"), p.nsrc, p.nst)
    assert np.array_equal(xyz, xyz2) and np.array_equal(stf, stf2)
    pts = awpsgt.read_cordfile(os.path.join(d, "cordfile"))
    assert np.array_equal(pts, orc.read_cordfile(os.path.join(d, "cordfile")))


@pytest.mark.parametrize("name", list(cases.CASES))
def test_sgt_point_rule_matches_oracle(case_dirs, name):
    d, _ = case_dirs[name]
    p = awpsgt.read_in3d(os.path.join(d, "IN3D"))
    pts = awpsgt.read_cordfile(os.path.join(d, "cordfile"))
    keep, ijk = awpsgt.sgt_select(p, pts)
    q = orc.read_in3d(os.path.join(d, "IN3D"))
    media = orc.read_media(os.path.join(d, "media"), q["NX"], q["NY"], q["NZ"], q["NVAR"])
    xyz, stf = orc.read_source(os.path.join(d, "source"), q["NSRC"], q["NST"])
    o = orc.Oracle(q, media, xyz, stf, pts)
    kidx, kpos = o.points()
    assert np.array_equal(np.nonzero(keep)[0], kidx)
    assert np.array_equal(ijk[keep.astype(bool)], kpos)


def test_sgt_point_rule_brute_force():
    """Enumerate the NPX x NPY x NPZ ranks as station.f:31-58 does, for every
    point of a box around a small grid, and compare with the C rule."""
    p = awpsgt.In3d()
    p.nx, p.ny, p.nz, p.npx, p.npy, p.npz = 12, 10, 8, 3, 2, 2
    nxt, nyt, nzt = 4, 5, 4
    pts = np.array([[x, y, z] for x in range(0, 15) for y in range(0, 12) for z in range(-1, 10)],
                   dtype=np.int32)
    keep, ijk = awpsgt.sgt_select(p, pts)
    for n, (x, y, z) in enumerate(pts):
        owner = None
        for cx in range(3):
            for cy in range(2):
                for cz in range(2):
                    if (int((p.nz - z) / nzt) == cz and int((y - 1) / nyt) == cy
                            and int((x - 1) / nxt) == cx):
                        a = x % nxt; lx = a if a else nxt
                        a = y % nyt; ly = a if a else nyt
                        a = (p.nz - z + 1) % nzt; lz = a if a else nzt
                        owner = (cx * nxt + lx, cy * nyt + ly, cz * nzt + lz)
        if owner is None:
            assert keep[n] == 0, (x, y, z)
        else:
            assert keep[n] == 1 and tuple(ijk[n]) == owner, (x, y, z)
    # in-range points map to (x, y, nz+1-z)
    inr = (pts[:, 0] >= 1) & (pts[:, 0] <= 12) & (pts[:, 1] >= 1) & (pts[:, 1] <= 10) & \
          (pts[:, 2] >= 1) & (pts[:, 2] <= 8)
    assert np.all(keep[inr] == 1)
    assert np.array_equal(ijk[inr], np.stack([pts[inr, 0], pts[inr, 1], 9 - pts[inr, 2]], 1))


def test_sgt_point_rule_rejects_undefined():
    p = awpsgt.In3d()
    p.nx, p.ny, p.nz, p.npx, p.npy, p.npz = 12, 10, 8, 1, 1, 1
    for bad in ([-1, 3, 3], [3, -2, 3], [3, 3, 10]):
        with pytest.raises(awpsgt.AwpError):
            awpsgt.sgt_select(p, np.array([bad], dtype=np.int32))


@pytest.mark.parametrize("nt,ntiskp,ws", [(120, 10, 12), (133, 10, 5), (1500, 10, 150),
                                          (1500, 10, 100), (99, 10, 12), (4001, 10, 400)])
def test_sgt_file_geometry(nt, ntiskp, ws):
    p = awpsgt.In3d()
    p.nt, p.ntiskp_sgt, p.write_step = nt, ntiskp, ws
    nt_out, nflush = awpsgt.sgt_file_geometry(p)
    nt_out2, flushes = orc.sgt_file_layout(nt, ntiskp, ws)
    assert nt_out == nt_out2 and nflush == len(flushes)


def test_golden_file_sizes_follow_geometry():
    """The reference's file sizes in the fixtures equal the closed form."""
    for name in cases.CASES:
        spec, media, pts = cases.tiny_inputs(name)
        nt_out, flushes = orc.sgt_file_layout(spec.nt, spec.ntiskp_sgt, spec.ws())
        g = load_golden(name)
        p = awpsgt.In3d()
        p.nx, p.ny, p.nz, p.npx, p.npy, p.npz = spec.nx, spec.ny, spec.nz, spec.npx, spec.npy, spec.npz
        keep, _ = awpsgt.sgt_select(p, pts.astype(np.int32))
        last = int(np.nonzero(keep)[0][-1])
        size = last * 24 * nt_out + 5 * 4 * nt_out + (len(flushes) - 1) * 4 * spec.ws() + 4 * spec.ws()
        assert int(g["nbytes"]) == size, name


def _write(tmp_path, **over):
    path = os.path.join(str(tmp_path), "IN3D")
    base = {"NX": 64, "NY": 64, "NZ": 32, "ND": 8, "TMAX": 0.99, "DT": 0.01}
    base.update(over)
    write_in3d(path, base)
    return path


@pytest.mark.parametrize("over,msg", [
    ({"IGREEN": 0}, "igreen"), ({"IGREEN": -1}, "igreen"), ({"NPC": 1}, "NPC"),
    ({"NVE": 0}, "NVE"), ({"IFAULT": 2}, "IFAULT"), ({"MEDIARESTART": 0}, "MEDIARESTART"),
    ({"NVAR": 4}, "NVAR"), ({"ND": 40}, "ND"), ({"NPX": 3}, "NPX"), ({"ARBC": 1.5}, "ARBC"),
    ({"NTISKP_SGT": 0}, "NTISKP_SGT"),
])
def test_unsupported_settings_rejected(tmp_path, over, msg):
    p = awpsgt.read_in3d(_write(tmp_path, **over))
    with pytest.raises(awpsgt.AwpError, match=msg):
        awpsgt.validate_in3d(p)


def test_in3d_list_directed_forms(tmp_path):
    """Blank lines, trailing labels, d-exponents, quoted names with blanks."""
    path = _write(tmp_path)
    text = open(path).read().replace("     0.01 DT", "    1.0d-2 DT  = time step")
    text = text.replace("'sgt' SGTGRO", "'out dir/sgt file' SGTGRO")
    text = "\n\n" + text
    open(path, "w").write(text)
    p = awpsgt.read_in3d(path)
    assert np.float32(p.dt) == np.float32(0.01)
    assert p.s("sgtgro") == "out dir/sgt file"
    assert p.nt == 100


def test_in3d_truncated_rejected(tmp_path):
    path = _write(tmp_path)
    lines = [ln for ln in open(path).read().splitlines() if ln.strip()]
    open(path, "w").write("\n".join(lines[:50]) + "\n")
    with pytest.raises(awpsgt.AwpError, match="records"):
        awpsgt.read_in3d(path)


def test_source_reader_errors(tmp_path):
    src = os.path.join(str(tmp_path), "src")
    open(src, "w").write("3 4 1\n1.0, 0, 0, 0, 0, 0\n")
    with pytest.raises(awpsgt.AwpError, match="sample 2"):
        awpsgt.read_source(src, 1, 2)
    xyz, stf = awpsgt.read_source(src, 1, 1)
    assert list(xyz[0]) == [3, 4, 1] and stf[0, 0, 0] == 1.0


def test_cordfile_comments(tmp_path):
    f = os.path.join(str(tmp_path), "cord")
    open(f, "w").write("# header\n# more\n3\n1 2 3\n4 5 6\n7 8 9\n")
    assert awpsgt.read_cordfile(f).tolist() == [[1, 2, 3], [4, 5, 6], [7, 8, 9]]
