"""Input builders (awpsgt.builders) against the reference tools: the mesh
reformatter built from its C source (oracle/_ref/reformat_velocity), and
closed-form checks of the point-list, source and IN3D builders."""
import os
import subprocess

import numpy as np
import pytest

from conftest import REPO
import awpsgt
from awpsgt import builders as B

TOOL = os.path.join(REPO, "oracle", "_ref", "reformat_velocity")


@pytest.mark.skipif(not os.path.exists(TOOL), reason="oracle/_ref/reformat_velocity not built")
def test_velocity_reformat_matches_reference_tool(tmp_path):
    nx, ny, nz = 7, 5, 4            # RWG grid (x fastest, then z, then y)
    rng = np.random.default_rng(3)
    pre = os.path.join(str(tmp_path), "mesh")
    arrs = []
    for ext in ("p", "s", "d"):
        a = rng.uniform(0.3, 8.0, size=(ny, nz, nx)).astype(np.float32)
        a.tofile(pre + "." + ext)
        arrs.append(a)
    out_ref = os.path.join(str(tmp_path), "ref")
    subprocess.run([TOOL, str(nx), str(ny), str(nz), pre, out_ref], check=True, stdout=subprocess.DEVNULL)
    ref = open(out_ref, "rb").read()
    mine = B.reformat_velocity_arrays(*arrs)
    assert mine.shape == (nz, nx, ny, 3)
    assert mine.tobytes() == ref
    out = os.path.join(str(tmp_path), "mine")
    B.reformat_velocity(nx, ny, nz, pre, out, chunk_y=2)     # banded file path
    assert open(out, "rb").read() == ref
    # AWP x is RWG y: the value at AWP (x, y, z) = RWG (y, x, z) * 1000
    m = np.frombuffer(ref, np.float32).reshape(nz, nx, ny, 3)
    assert m[2, 4, 1, 1] == np.float32(np.float64(arrs[1][1, 2, 4]) * 1000.0)


def test_cordfile_swap_dedupe_sort(tmp_path):
    rwg = ["hdr\n", "hdr\n", "hdr\n", "hdr\n", "4\n", "3 1 5\n", "0 2 1 extra\n", "3 1 5\n", "1 0 2\n"]
    src, dst = os.path.join(str(tmp_path), "r"), os.path.join(str(tmp_path), "a")
    open(src, "w").writelines(rwg)
    rc = B.gen_awp_cordfile(src, dst)
    assert rc == 2                                # 4 declared, 3 unique
    # (y+1, x+1, z), sorted by AWP y (= RWG x + 1), then AWP x, then z
    assert open(dst).read() == "3\n3 1 1 \n1 2 2 \n2 4 5 \n"
    pts, n_in, dups = B.rwg_to_awp_points(rwg)
    assert n_in == 4 and dups == ["Duplicate point entry 2 4 5"]
    # the output is a cordfile the solver reads (x y z per line after the count)
    pts_read = awpsgt.read_cordfile(dst)
    assert pts_read.tolist() == [list(p) for p in pts]


def test_source_file_swaps_and_offsets(tmp_path):
    text = B.awp_source_text("12 34\n", ["1.0 0 0 0 0 0\n", "2.0 0 0 0 0 0\n"])
    assert text == "35 13 1\n1.0 0 0 0 0 0\n2.0 0 0 0 0 0\n"


def _gridout(nx_rwg, ny_rwg, nz):
    lines = ["xlen=1\n", "nx=%d\n" % nx_rwg] + ["%d\n" % i for i in range(nx_rwg + 1)]
    lines += ["ny=%d\n" % ny_rwg] + ["%d\n" % i for i in range(ny_rwg + 1)]
    lines += ["nz=%d\n" % nz] + ["%d\n" % i for i in range(nz + 1)]
    return lines


def test_in3d_builder_reads_back(tmp_path):
    dims = B.gridout_dims(_gridout(200, 300, 100))  # RWG x = 200 -> AWP NY
    assert dims == (300, 200, 100)
    P = B.in3d_params("SITE", dims, "y", 0.5, (2, 2, 1), "awp.mesh")
    assert P["igreen"] == 2 and P["DH"] == 200.0 and P["DT"] == 0.01
    assert P["NST"] == 20000 and P["WRITE_STEP"] == 2000      # 200 s / 0.01, a multiple of 1000
    path = os.path.join(str(tmp_path), "IN3D")
    open(path, "w").write(B.in3d_text(P, "y"))
    p = awpsgt.read_in3d(path)
    assert (p.nx, p.ny, p.nz, p.igreen, p.nst, p.npx, p.npy, p.npz) == (300, 200, 100, 2, 20000, 2, 2, 1)
    assert p.s("sgtgro") == "comp_y/output_sgt/awp-strain-SITE-fy"
    assert p.s("insgt") == "comp_y/input/awp.SITE.cordfile"
    assert p.ntiskp == 40000 and p.ntiskp_sgt == 10 and p.mediarestart == 2
    assert abs(p.tmax - 200.0) < 1e-3 and p.nt == 20001
    awpsgt.validate_in3d(p)
    with pytest.raises(ValueError, match="factor"):
        B.in3d_params("S", dims, "x", 0.5, (7, 1, 1), "m")
    P3 = B.in3d_params("S", dims, "x", 1.0, (1, 1, 1), "m", ref=dict(B.IN3D_REF, TMAX="100.3"))
    assert P3["DH"] == 100.0 and P3["NST"] == 21000   # int(100.3 / 0.005) = 20060 -> 21000
