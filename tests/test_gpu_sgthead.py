"""The device's RWG-order SGT writer (awp_write_sgt_file_rwg) against SgtHead's
reformatter applied to the device's own raw file: the numpy restatement
(pinned to the reference tool in test_sgthead.py) and, where it was built,
the reference tool itself (oracle/_ref/reformat_awp)."""
import os
import subprocess

import numpy as np

import pytest

from conftest import REPO
import awpsgt
from oracle.sgthead import reformat_awp
from test_gpu_parity import _gpu_run, _inputs

pytestmark = pytest.mark.gpu

TOOL = os.path.join(REPO, "oracle", "_ref", "reformat_awp")


@pytest.mark.parametrize("name", ["g1_fx_layered", "g3_fz_surface", "g7_random_edges"])
@pytest.mark.parametrize("z", [False, True])
def test_rwg_file_equals_reformatted_raw_file(case_dirs, name, z, tmp_path):
    d, _ = case_dirs[name]
    p, media, xyz, stf, pts = _inputs(d)
    raw, rwg = os.path.join(str(tmp_path), "raw"), os.path.join(str(tmp_path), "rwg")
    with _gpu_run(p, media, xyz, stf, pts) as r:
        r.write_sgt_file(raw)
        r.write_sgt_file_rwg(rwg, z)
    nt_out, _ = awpsgt.sgt_file_geometry(p)
    mine = open(rwg, "rb").read()
    assert mine == reformat_awp(open(raw, "rb").read(), nt_out, z)
    assert len(mine) > 0 and len(mine) % (24 * nt_out) == 0
    if os.path.exists(TOOL):
        out = os.path.join(str(tmp_path), "tool")
        subprocess.run([TOOL, raw, str(nt_out), out] + (["-z"] if z else []), check=True,
                       stdout=subprocess.DEVNULL)
        assert mine == open(out, "rb").read()


def test_media_from_rwg_mesh_equals_reformatted_file(case_dirs):
    """awp_set_media_rwg (device reformat of the RWG mesh) gives the media
    awp_set_media gives from the file SgtHead's reformat_velocity writes
    (its restatement is pinned to the reference tool in test_builders.py)."""
    from awpsgt.builders import reformat_velocity_arrays
    d, _ = case_dirs["g1_fx_layered"]
    p, media, xyz, stf, pts = _inputs(d)
    assert p.nvar == 3
    rng = np.random.default_rng(8)
    shape = (p.nx, p.nz, p.ny)             # RWG [y][z][x] = AWP [x][z][y]
    vs = rng.uniform(0.4, 3.8, size=shape).astype(np.float32)
    vp = (vs * rng.uniform(1.6, 2.0, size=shape)).astype(np.float32)
    rho = rng.uniform(1.8, 2.9, size=shape).astype(np.float32)
    awp = reformat_velocity_arrays(vp, vs, rho)      # [nz][AWP y][AWP x][3]
    r1 = awpsgt.Run(p)
    r1.set_media_rwg(vp, vs, rho)
    r2 = awpsgt.Run(p)
    r2.set_media(awp)
    for f in ("d1", "mu", "lam", "qp", "qs"):
        assert np.array_equal(r1.field(f), r2.field(f)), f
    r1.close()
    r2.close()
