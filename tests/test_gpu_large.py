"""The HIP path at the workload sizes: the 1500x1000x300 site shape (450M
points), the 8-run batch of 400x300x150 .. 900x700x220 grids and the 16-run
queue of 500x400x150 .. 1200x900x280 grids (2.2G points in one batch).

At these sizes the CPU restatement finishes two steps of one grid in about a
minute, so the oracle pins the first steps bit for bit; the rest is held by
properties that do not depend on size: the fused one-sweep step, after the
tuner has chosen its tiles, z chunks and buffer placement, equals the
two-stage step bit for bit; two runs are bit-identical; doubling the source
doubles the SGT; every run of a batch equals the same run stepped alone.
"""
import os

import numpy as np
import pytest

import awpsgt
from awpsgt.batchq import Need, admit
from awpsgt.workloads import CONFIGS, batch_jobs, batch_run, c2_run
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

STATE = ["u1", "v1", "w1", "xx", "yy", "zz", "xy", "xz", "yz", "r1", "r2", "r3", "r4", "r5", "r6"]
KEXT = {"u1": 2, "v1": 2, "w1": 2, "zz": 2, "xz": 2, "yz": 2}


def _say(msg):
    print("[large] " + msg, flush=True)


def _fields(r, names):
    return {f: r.field(f, KEXT.get(f, 0)) for f in names}


def _in3d_dict(p):
    return {"NX": p.nx, "NY": p.ny, "NZ": p.nz, "NPX": p.npx, "NPY": p.npy, "NPZ": p.npz, "DH": p.dh, "DT": p.dt,
            "NT": p.nt, "ND": p.nd, "ARBC": p.arbc, "SOCALQ": p.socalq, "FL": p.fl, "FH": p.fh, "FP": p.fp,
            "IGREEN": p.igreen, "NST": p.nst, "NTISKP_SGT": p.ntiskp_sgt, "WRITE_STEP": p.write_step,
            "NVAR": p.nvar}


def _oracle_check(run, media_host, xyz, stf, pts, nsteps, names):
    """The run after nsteps against the FP32 restatement, field by field."""
    o = orc.Oracle(_in3d_dict(run.p), media_host, xyz, stf, pts)
    o.run(nsteps, 16)
    for f in names:
        k = KEXT.get(f, 0)
        a, b = run.field(f, k), o.field(f, k)
        assert np.array_equal(a, b), (f, float(np.abs(a - b).max()))
    o.close()


@pytest.fixture(scope="module")
def c2_tuned():
    """The tuned default run of the site shape after 20 steps: its scheme,
    SGT and state fields (host copies)."""
    _say("c2 tuned run: set-up")
    r, nkept = c2_run(20)
    fused, ty, nc = r.tune(2)
    _say("c2 tuned: %s, 64 x %d tiles, %d z chunks" % ("fused" if fused else "two-stage", ty, nc))
    r.step(20)
    r.synchronize()
    out = {"scheme": (fused, ty, nc), "sgt": r.sgt(), "fields": _fields(r, STATE), "nkept": nkept}
    r.close()
    return out


def test_c2_tuned_step_equals_two_stage(c2_tuned, monkeypatch):
    assert c2_tuned["nkept"] == 1_500_000
    assert c2_tuned["scheme"][0], "the tuner keeps the fused step at the site shape"
    monkeypatch.setenv("AWP_SCHEME", "split")
    _say("c2 two-stage run")
    r, _ = c2_run(20)
    assert r.scheme()[0] is False
    r.step(20)
    r.synchronize()
    s = r.sgt()
    assert np.array_equal(s, c2_tuned["sgt"])
    assert np.all(s[:, :, :2] != -1) and np.any(s[:, :, :2] != 0) and np.all(np.isfinite(s))
    for f in STATE:
        assert np.array_equal(r.field(f, KEXT.get(f, 0)), c2_tuned["fields"][f]), f
    r.close()


def test_c2_untuned_repeat_and_linearity(c2_tuned):
    """The default fused launch (untuned: another tile shape, chunking and
    buffer placement) gives the same bits; twice the source, twice the SGT."""
    _say("c2 untuned repeat")
    r, _ = c2_run(20)
    r.step(20)
    r.synchronize()
    assert np.array_equal(r.sgt(), c2_tuned["sgt"])
    assert np.array_equal(r.field("xz", 2), c2_tuned["fields"]["xz"])
    r.close()
    _say("c2 doubled source")
    r2, _ = c2_run(20, amp=2e9)
    r2.step(20)
    r2.synchronize()
    s1, s2 = c2_tuned["sgt"].astype(np.float64), r2.sgt().astype(np.float64)
    live = s1[:, :, :2]
    num = np.linalg.norm(s2[:, :, :2] - 2 * live)
    assert num <= 1e-6 * np.linalg.norm(2 * live), num
    r2.close()


@pytest.mark.timeout(900)
def test_c2_equals_restatement_after_two_steps():
    """Every state field of the site shape after 2 steps bit for bit against
    the FP32 CPU restatement (16 threads)."""
    from awpsgt.media import c2_media
    from awpsgt.workloads import build_params, sgt_points, source_of
    cfg = CONFIGS["c2"]
    p = build_params(cfg, 2)
    _say("c2 media for the restatement")
    m = c2_media(p.nx, p.ny, p.nz, p.dh, seed=cfg["seed"])
    media_host = m.cpu().numpy()
    del m
    r, _ = c2_run(2)
    r.step(2)
    r.synchronize()
    xyz, stf = source_of(p, cfg["src"])
    _say("c2 restatement, 2 steps")
    _oracle_check(r, media_host, xyz, stf, sgt_points(p, cfg["nsgt"], 3), 2, STATE)
    r.close()


def _batch_vs_single(config, idx, storage, nt, tune=True, budget_bytes=None):
    """Every run of the workload in one batch (one wave), then the runs
    `idx` alone: SGT and fields bit for bit."""
    import torch
    jobs = batch_jobs(config, nt)
    for j in jobs:
        j["need"] = Need(*awpsgt.estimate_bytes(j["p"], storage, CONFIGS[config]["nsgt"]))
    free, _ = torch.cuda.mem_get_info()
    waves = admit([j["need"] for j in jobs], budget_bytes or free - (8 << 30))
    assert len(waves) == 1, "the whole queue is admitted in one wave"
    _say("%s: set-up of %d runs" % (config, len(jobs)))
    runs = [batch_run(config, j, k, storage)[0] for k, j in enumerate(jobs)]
    with awpsgt.Batch(runs) as b:
        if tune:
            b.tune(1)
        _say("%s: batch %s, launch %s; %d steps" % (config, "fused" if b.fused() else "two-stage", b.launch(), nt))
        b.step(nt)
    got = {k: (runs[k].sgt(), _fields(runs[k], ["u1", "xx", "xz", "r1"])) for k in idx}
    for r in runs:
        r.close()
    torch.cuda.empty_cache()
    for k in idx:
        _say("%s: run %d (%dx%dx%d) alone" % (config, k, jobs[k]["p"].nx, jobs[k]["p"].ny, jobs[k]["p"].nz))
        r, _ = batch_run(config, jobs[k], k, storage)
        r.step(nt)
        r.synchronize()
        s, f = got[k]
        assert np.array_equal(r.sgt(), s), k
        assert np.any(s != 0)
        for name, v in f.items():
            assert np.array_equal(r.field(name, KEXT.get(name, 0)), v), (k, name)
        r.close()


@pytest.mark.timeout(900)
def test_c3_tuned_batch_equals_single_runs():
    """The 8 runs of 400x300x150 .. 900x700x220 (both wavefield buffers, so
    the tuner chooses between the fused and the two-stage batch step)."""
    _batch_vs_single("c3", list(range(8)), awpsgt.AWP_STORAGE_FP32, 20)


@pytest.mark.timeout(900)
def test_c3_largest_run_equals_restatement():
    from awpsgt.media import c2_media
    from awpsgt.workloads import sgt_points, source_of
    job = batch_jobs("c3", 2)[7]
    p = job["p"]
    assert (p.nx, p.ny, p.nz) == (900, 700, 220)
    m = c2_media(p.nx, p.ny, p.nz, p.dh, seed=job["seed"])
    media_host = m.cpu().numpy()
    del m
    r, _ = batch_run("c3", job, 7, awpsgt.AWP_STORAGE_FP32)
    r.step(2)
    r.synchronize()
    xyz, stf = source_of(p, (p.nx // 2, p.ny // 2, 10))
    _say("c3 largest run: restatement, 2 steps")
    _oracle_check(r, media_host, xyz, stf, sgt_points(p, CONFIGS["c3"]["nsgt"], 107), 2, STATE)
    r.close()


@pytest.mark.timeout(900)
def test_c5_queue_in_one_batch_smallest_and_largest(monkeypatch):
    """All 16 runs of the mixed-storage queue (500x400x150 .. 1200x900x280,
    2.2G points, one wave under the device budget) in one batch: the
    smallest and the largest run each equal their single run bit for bit."""
    monkeypatch.setenv("AWP_SCHEME", "split")
    _batch_vs_single("c5", [0, 15], awpsgt.AWP_STORAGE_MIXED, 10)


@pytest.mark.timeout(900)
@pytest.mark.parametrize("storage,kc", [(awpsgt.AWP_STORAGE_FP32, 0), (awpsgt.AWP_STORAGE_FP32, 200),
                                        (awpsgt.AWP_STORAGE_MIXED, 0)])
def test_arrays_beyond_4gib_fused_equals_two_stage(storage, kc, monkeypatch):
    """2000x1500x400 (1.2G points: every volume array is 4.98 GB, beyond
    4 GiB): the fused sweep (32-bit offsets from a base rebased per z chunk;
    default chunks, or 2 chunks of 200 levels) equals the two-stage step,
    which addresses these arrays with 64-bit element indices, bit for bit
    after 3 steps."""
    from awpsgt.media import c2_media
    from awpsgt.workloads import build_params, make_run
    cfg = CONFIGS["c2"]
    p = build_params(cfg, 3, 2000, 1500, 400)
    names = ["u1", "w1", "xx", "xz", "yz", "r1", "r5"]

    def run(scheme):
        monkeypatch.setenv("AWP_SCHEME", scheme)
        if kc and scheme == "fused":
            monkeypatch.setenv("AWP_KC", str(kc))
        else:
            monkeypatch.delenv("AWP_KC", raising=False)
        _say("4 GiB arrays: %s run, storage %d" % (scheme, storage))
        r, _ = make_run(p, storage, lambda: c2_media(p.nx, p.ny, p.nz, p.dh, seed=5), (1000, 750, 10), 10000, 9)
        fused, ty, nc = r.scheme()
        assert fused == (scheme == "fused"), (scheme, fused)
        r.step(3)
        r.synchronize()
        out = (r.sgt(), _fields(r, names), nc)
        r.close()
        import torch
        torch.cuda.empty_cache()
        return out

    sf, ff, nc = run("fused")
    if kc:
        assert nc == 2
    ss, fs, _ = run("split")
    assert np.array_equal(sf, ss)
    for f in names:
        assert np.array_equal(ff[f], fs[f]), f
    assert np.any(ff["u1"] != 0)
