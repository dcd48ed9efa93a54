"""awp_get_field_noalloc (readback through the library's fixed device scratch,
no device allocation) against awp_get_field, bit for bit, on two calls in a
row -- a 1500x1000x12 run whose fields take three chunks of levels, then a
small run -- in both storage modes, and its rejection of a grid whose level
exceeds the scratch."""
import numpy as np
import pytest

import awpsgt
from awpsgt.workloads import CONFIGS, build_params, layered_nvar5, make_run

pytestmark = pytest.mark.gpu

NAMES = ["u1", "w1", "xx", "xz", "r1", "r6", "d1", "lam", "qs"]


def _check(r, kext_names=("u1", "xz")):
    for f in NAMES:
        a = r.field(f)
        b = r.field_noalloc(f)
        assert np.array_equal(a, b), f
    for f in kext_names:
        assert np.array_equal(r.field(f, 2), r.field_noalloc(f, 2)), f


@pytest.mark.parametrize("storage", [awpsgt.AWP_STORAGE_FP32, awpsgt.AWP_STORAGE_MIXED])
def test_noalloc_readback_equals_field(storage):
    big = build_params(CONFIGS["c2"], 6, 1500, 1000, 48)
    big.nz, big.nd = 12, 4             # 1.5M-point levels: 5 per chunk of the scratch
    awpsgt.validate_in3d(big)
    r, _ = make_run(big, storage, lambda: layered_nvar5(big), (750, 500, 6), 2000, 4)
    r.step(6)
    r.synchronize()
    print("[noalloc] 1500x1000x12 storage %d" % storage, flush=True)
    _check(r)
    assert np.any(r.field_noalloc("u1") != 0)
    r.close()
    small = build_params(CONFIGS["c2"], 5, 96, 64, 48)
    r, _ = make_run(small, storage, lambda: layered_nvar5(small), (48, 32, 5), 200, 5)
    r.step(5)
    r.synchronize()
    print("[noalloc] 96x64x48 storage %d" % storage, flush=True)
    _check(r)
    r.close()


def test_noalloc_rejects_levels_beyond_the_scratch():
    p = build_params(CONFIGS["c2"], 2, 3000, 3000, 48)
    p.nz, p.nd = 8, 2                 # 3000 x 3000 levels (9M points > 8M), few of them
    awpsgt.validate_in3d(p)
    r = awpsgt.Run(p)
    with pytest.raises(Exception, match="exceeds"):
        r.field_noalloc("u1")
    r.close()
