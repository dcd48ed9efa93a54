"""The device's RWG-order SGT writer (awp_write_sgt_file_rwg) against SgtHead's
reformatter applied to the device's own raw file: the numpy restatement
(pinned to the reference tool in test_sgthead.py) and, where it was built,
the reference tool itself (oracle/_ref/reformat_awp)."""
import os
import subprocess

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
