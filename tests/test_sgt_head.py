"""The SGT header writer (awp_write_sgt_head, SgtHead write_head's output)
against the reference tool's files for a seeded case (tests/golden/
sgt_head.npz, written by tests/golden/make_head_golden.py from
oracle/_ref/write_head), against the tool itself on other media where it is
built, and against what the geometry fixes (the model-box centre projects to
the box's geographic origin; distances; harmonic density)."""
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

from conftest import REPO
import awpsgt

sys.path.insert(0, os.path.join(REPO, "tests", "golden"))
import head_case as hc  # noqa: E402

TOOL = os.path.join(REPO, "oracle", "_ref", "write_head")
needs_tool = pytest.mark.skipif(not os.path.exists(TOOL), reason="oracle/_ref/write_head not built")

MASTER = struct.Struct("<i5f3i")           # sgtmaster, 36 B
INDEX = struct.Struct("<q3if")             # sgtindex, 24 B
HEADER = struct.Struct("<qi5fi7f3i3f3i7f")  # sgtheader, 128 B
HFIELDS = ("indx geoproj modellon modellat modelrot xshift yshift nt xazim dt tst h src_lat src_lon src_dep "
           "xsrc ysrc zsrc sgt_lat sgt_lon sgt_dep xsgt ysgt zsgt cdist lam mu rho xmom ymom zmom").split()


def _mine(p, comp, corr, out):
    a = hc.ARGS
    awpsgt.write_sgt_head(p["modelbox"], p["coordfile"], p["fdloc"], p["gridout"], a["spacing"], a["nt"], a["dt"],
                          a["sgt_tinc"], comp, a["moment"], a["flo"], p["media"], out, mu_corr=corr)
    return open(out, "rb").read()


def _parse(raw):
    m = MASTER.unpack_from(raw, 0)
    n = m[6]
    idx = [INDEX.unpack_from(raw, MASTER.size + k * INDEX.size) for k in range(n)]
    off = MASTER.size + n * INDEX.size
    hdr = [dict(zip(HFIELDS, HEADER.unpack_from(raw, off + k * HEADER.size))) for k in range(n)]
    assert len(raw) == off + n * HEADER.size
    return m, idx, hdr


@pytest.mark.parametrize("comp", ["x", "z"])
@pytest.mark.parametrize("corr", [False, True])
def test_header_equals_reference_file(comp, corr, tmp_path):
    p = hc.write_inputs(str(tmp_path))
    mine = _mine(p, comp, corr, os.path.join(str(tmp_path), "out"))
    want = np.load(os.path.join(REPO, "tests", "golden", "sgt_head.npz"))["%s_%s" % (comp, "corr" if corr else "plain")]
    assert mine == want.tobytes()


@needs_tool
@pytest.mark.parametrize("seed", [11, 12])
def test_header_equals_tool_on_other_media(seed, tmp_path):
    import make_head_golden as mk
    p = hc.write_inputs(str(tmp_path), seed=seed)
    for comp, corr in (("y", False), ("x", True)):
        want = mk.run_tool(p, comp, corr, os.path.join(str(tmp_path), "ref"))
        assert _mine(p, comp, corr, os.path.join(str(tmp_path), "mine")) == want, (seed, comp, corr)


def test_header_geometry_and_media(tmp_path):
    p = hc.write_inputs(str(tmp_path))
    m, idx, hdr = _parse(_mine(p, "y", False, os.path.join(str(tmp_path), "out")))
    assert m[0] == 1 and m[6] == m[7] == 40 and m[8] == hc.ARGS["nt"] // hc.ARGS["sgt_tinc"]
    h = hc.SPACING
    assert m[4] == pytest.approx(-0.5 * hc.NX * h) and m[5] == pytest.approx(-0.5 * hc.NY * h)
    media = np.fromfile(p["media"], dtype="<f4").reshape(hc.NZ, hc.NX, hc.NY, 3).astype(np.float64) * 1e-3
    for r in hdr:
        # the site sits at the model-box centre: it projects onto the box's origin
        assert r["src_lon"] == pytest.approx(-118.286, abs=2e-5)
        assert r["src_lat"] == pytest.approx(34.0192, abs=2e-5)
        assert r["ymom"] == pytest.approx(1e20) and r["xmom"] == 0.0 and r["zmom"] == 0.0
        assert r["tst"] == pytest.approx(-3.0) and r["dt"] == pytest.approx(0.05) and r["xazim"] == pytest.approx(35.0)
        x, y, z = r["xsgt"], r["ysgt"], r["zsgt"]
        d = np.sqrt((x - 10.0) ** 2 + (y - 8.0) ** 2 + (z - 1.0) ** 2) * h
        assert r["cdist"] == pytest.approx(d, rel=1e-6)
        # moduli from the plane above the point (its own plane at the surface),
        # density from y = 1 on the y = 0 row
        pa = max(z - 1, 0)
        vp, vs = media[pa, x, y, 0], media[pa, x, y, 1]
        rho0 = media[pa, x, max(y, 1), 2]
        assert r["mu"] == pytest.approx(1e10 * vs * vs * rho0, rel=2e-6)
        assert r["lam"] == pytest.approx(1e10 * (vp * vp * rho0 - 2 * vs * vs * rho0), rel=2e-5)
        if x == hc.NX - 1 or z == hc.NZ - 1:
            assert r["rho"] == pytest.approx(rho0, rel=1e-6)
        else:
            nb = [media[pa, x + 1, max(y, 1), 2], media[pa, x, y + 1, 2] if y + 1 < hc.NY else media[pa, x + 1, 0, 2],
                  media[pa + 1, x, max(y, 1), 2]]
            want = 3.0 / sum(2.0 / (rho0 + q) for q in nb)
            assert r["rho"] == pytest.approx(want, rel=2e-6)


def test_header_rejects_point_outside_grid(tmp_path):
    p = hc.write_inputs(str(tmp_path))
    txt = open(p["coordfile"]).read().splitlines()
    txt[3] = "%d 0 0 1 0 0 0" % hc.NX
    open(p["coordfile"], "w").write("\n".join(txt) + "\n")
    with pytest.raises(awpsgt.AwpError, match="outside"):
        _mine(p, "x", False, os.path.join(str(tmp_path), "out"))
