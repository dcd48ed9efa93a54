import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (REPO, os.path.join(REPO, "cybershake-core_amd"), os.path.join(REPO, "tests", "golden")):
    if p not in sys.path:
        sys.path.insert(0, p)

import cases  # noqa: E402

GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X)")
    config.addinivalue_line("markers", "slow: long CPU run, enabled by AWP_SLOW=1")


def pytest_collection_modifyitems(config, items):
    if os.environ.get("AWP_SLOW") == "1":
        return
    skip = pytest.mark.skip(reason="slow CPU check; set AWP_SLOW=1")
    for it in items:
        if "slow" in it.keywords:
            it.add_marker(skip)


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def case_dirs(tmp_path_factory):
    """Run directories (IN3D, media, source, cordfile) of every seeded case,
    written once per session; maps name -> (dir, inputs_sha256)."""
    out = {}
    root = tmp_path_factory.mktemp("cases")
    for name in cases.CASES:
        spec, media, pts = cases.tiny_inputs(name)
        d = os.path.join(str(root), name)
        out[name] = (d, cases.write_case(d, spec, media, pts))
    return out


@pytest.fixture(scope="session")
def c1_dir(tmp_path_factory):
    spec, media, pts = cases.c1_inputs()
    d = os.path.join(str(tmp_path_factory.mktemp("c1")), "c1")
    return d, cases.write_case(d, spec, media, pts)


@pytest.fixture(scope="session")
def c4_dir(tmp_path_factory):
    """800x600x200, 1500 steps, 2000 SGT points (the mixed-storage error case)."""
    spec, media, pts = cases.c4_inputs()
    d = os.path.join(str(tmp_path_factory.mktemp("c4")), "c4")
    sha = cases.write_case(d, spec, media, pts)
    del media
    return d, sha


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a - b))
