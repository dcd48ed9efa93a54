"""The SgtHead reformatter restatement against the reference tool built from
its own source (oracle/_ref/reformat_awp), on the reference solver's raw SGT
files (including dropped-point holes) and on hand-made records that pin the
sign of zero."""
import os
import subprocess

import numpy as np
import pytest

from conftest import REPO, load_golden
import cases
from oracle import oracle as orc
from oracle.sgthead import reformat_awp

TOOL = os.path.join(REPO, "oracle", "_ref", "reformat_awp")
needs_tool = pytest.mark.skipif(not os.path.exists(TOOL), reason="oracle/_ref/reformat_awp not built")


def _run_tool(raw: bytes, timesteps: int, z: bool, tmp_path) -> bytes:
    src, dst = os.path.join(str(tmp_path), "in"), os.path.join(str(tmp_path), "out")
    open(src, "wb").write(raw)
    cmd = [TOOL, src, str(timesteps), dst] + (["-z"] if z else [])
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
    return open(dst, "rb").read()


@needs_tool
@pytest.mark.parametrize("name", list(cases.CASES))
@pytest.mark.parametrize("z", [False, True])
def test_reformat_restatement_matches_tool(name, z, tmp_path):
    g = load_golden(name)
    raw = np.asarray(g["sgt"], dtype="<f4").tobytes()
    spec = cases.tiny_spec(name)
    nt_out, _ = orc.sgt_file_layout(spec.nt, spec.ntiskp_sgt, spec.ws())
    assert _run_tool(raw, nt_out, z, tmp_path) == reformat_awp(raw, nt_out, z)


def test_reformat_closed_form():
    ts = 3
    rec = np.arange(1, 19, dtype=np.float32).reshape(6, ts)    # XX..YZ = 1..18
    rec[4, 0] = 0.0                                             # +0 in XZ
    raw = rec.tobytes() + np.zeros(ts * 2, np.float32).tobytes()   # + an incomplete record
    out = np.frombuffer(reformat_awp(raw, ts), dtype="<f4").reshape(-1, 6, ts)
    assert out.shape[0] == 1
    assert np.array_equal(out[0, 0], rec[1]) and np.array_equal(out[0, 1], rec[0])
    assert np.array_equal(out[0, 2], rec[2]) and np.array_equal(out[0, 3], rec[3])
    assert np.array_equal(out[0, 4], -rec[5]) and np.array_equal(out[0, 5], -rec[4])
    assert np.signbit(out[0, 5, 0])                              # -(+0) = -0
    outz = np.frombuffer(reformat_awp(raw, ts, True), dtype="<f4").reshape(-1, 6, ts)
    assert np.array_equal(outz[0, 0], -2 * rec[1]) and np.array_equal(outz[0, 4], 2 * rec[5])
    assert not np.signbit(outz[0, 5, 0])                          # -(-2 * +0) = +0


def _fp16_file(path, scale, data, nt_out, ntiskp=10, dt=0.01):
    npts, _, nsamp = data.shape
    hdr = bytearray(64)
    hdr[0:8] = b"AWPSGT16"
    hdr[8:28] = np.array([1, npts, 6, nsamp, ntiskp], "<u4").tobytes()
    hdr[28:32] = np.array([dt], "<f4").tobytes()
    hdr[32:36] = np.array([nt_out], "<u4").tobytes()
    with open(path, "wb") as f:
        f.write(bytes(hdr))
        f.write(np.asarray(scale, "<f4").tobytes())
        f.write(np.asarray(data, "<f2").tobytes())


@pytest.mark.parametrize("layout", [0, 1, 2])
def test_fp16_point_reader_layouts(layout, tmp_path):
    """Decoding a range of points of the FP16 format gives the raw file's
    records (zero past the stored samples, zero for a dropped point) or, in
    the RWG layouts, those records as reformat_awp rewrites them."""
    import awpsgt
    rng = np.random.default_rng(5)
    npts, nsamp, nt_out = 6, 7, 9
    data = rng.uniform(-3e4, 3e4, size=(npts, 6, nsamp)).astype(np.float16)
    scale = np.ldexp(np.float32(1), rng.integers(-40, -20, size=(npts, 6))).astype(np.float32)
    scale[3] = 0.0                     # a point the rule dropped
    data[3] = 0
    path = os.path.join(str(tmp_path), "s16")
    _fp16_file(path, scale, data, nt_out)
    raw = np.zeros((npts, 6, nt_out), np.float32)
    raw[:, :, :nsamp] = data.astype(np.float32) * scale[:, :, None]
    want = raw if layout == 0 else np.frombuffer(
        reformat_awp(raw.tobytes(), nt_out, layout == 2), "<f4").reshape(npts, 6, nt_out)
    got = awpsgt.read_sgt16_points(path, 1, 4, nt_out, layout)
    assert got.tobytes() == want[1:5].tobytes()
    with pytest.raises(awpsgt.AwpError, match="beyond"):
        awpsgt.read_sgt16_points(path, 4, 3, nt_out, layout)
