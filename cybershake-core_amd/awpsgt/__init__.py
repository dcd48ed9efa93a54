"""Python front end of libawpsgt (include/awpsgt.h): the MI355X AWP-ODC-SGT
strain-Green-tensor solver.

The library is loaded from ``cybershake-core_amd/lib/libawpsgt.so`` (built by
``awpsgt.build.build()``).  There is no fallback: if the library is missing or
a HIP call fails, the call raises.

Typical use mirrors the reference's run of ``pmcl3d IN3D``
(AWP-ODC-SGT/src/pmcl3d.f):

    p = read_in3d("IN3D")
    run = Run(p)
    run.set_media(media)          # float32 [nz][ny][nx][nvar], file layout
    run.set_source(xyz, stf)      # [nsrc][3], [nsrc][nst][6]
    run.set_sgt_points(pts)       # [npts][3] cordfile order
    run.init(); run.step(p.nt)
    run.write_sgt_file(p.sgtgro)
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence, Tuple

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libawpsgt.so")

AWP_STORAGE_FP32 = 0
AWP_STORAGE_MIXED = 1

FIELDS = ["u1", "v1", "w1", "xx", "yy", "zz", "xy", "xz", "yz",
          "r1", "r2", "r3", "r4", "r5", "r6", "d1", "mu", "lam", "qp", "qs"]

_S = ctypes.c_char * 256


class In3d(ctypes.Structure):
    """Mirror of ``awp_in3d`` (include/awpsgt.h)."""
    _fields_ = [
        ("igreen", ctypes.c_int), ("tmax", ctypes.c_float), ("dh", ctypes.c_float),
        ("dt", ctypes.c_float), ("npc", ctypes.c_int), ("nd", ctypes.c_int),
        ("arbc", ctypes.c_float), ("pht", ctypes.c_float), ("nsrc", ctypes.c_int),
        ("nst", ctypes.c_int), ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("npx", ctypes.c_int), ("npy", ctypes.c_int), ("npz", ctypes.c_int),
        ("ifault", ctypes.c_int), ("checkpoint", ctypes.c_int), ("isfcvlm", ctypes.c_int),
        ("imd5", ctypes.c_int), ("ivelocity", ctypes.c_int), ("mediarestart", ctypes.c_int),
        ("nvar", ctypes.c_int), ("iost", ctypes.c_int), ("partdeg", ctypes.c_int),
        ("io_opt", ctypes.c_int), ("perf_meas", ctypes.c_int), ("idyna", ctypes.c_int),
        ("socalq", ctypes.c_int), ("nve", ctypes.c_int), ("mu_s", ctypes.c_float),
        ("mu_d", ctypes.c_float), ("fl", ctypes.c_float), ("fh", ctypes.c_float),
        ("fp", ctypes.c_float), ("read_step", ctypes.c_int), ("write_step", ctypes.c_int),
        ("write_step2", ctypes.c_int),
        ("nbgx", ctypes.c_int), ("nedx", ctypes.c_int), ("nskpx", ctypes.c_int),
        ("nbgy", ctypes.c_int), ("nedy", ctypes.c_int), ("nskpy", ctypes.c_int),
        ("nbgz", ctypes.c_int), ("nedz", ctypes.c_int), ("nskpz", ctypes.c_int),
        ("ntiskp", ctypes.c_int),
        ("nbgx2", ctypes.c_int), ("nedx2", ctypes.c_int), ("nskpx2", ctypes.c_int),
        ("nbgy2", ctypes.c_int), ("nedy2", ctypes.c_int), ("nskpy2", ctypes.c_int),
        ("nbgz2", ctypes.c_int), ("nedz2", ctypes.c_int), ("nskpz2", ctypes.c_int),
        ("ntiskp2", ctypes.c_int), ("ntiskp_sgt", ctypes.c_int),
        ("chkp", _S), ("chkj", _S), ("insrc", _S), ("invel", _S), ("insgt", _S),
        ("sxgro", _S), ("sygro", _S), ("szgro", _S), ("sxgro2", _S), ("sygro2", _S),
        ("szgro2", _S), ("sgtgro", _S), ("sgsn", _S),
        ("nt", ctypes.c_int),
    ]

    def s(self, name: str) -> str:
        return getattr(self, name).decode()


class HeadParams(ctypes.Structure):
    """awp_head_params (include/awpsgt.h)."""
    _fields_ = [("modelbox", ctypes.c_char_p), ("coordfile", ctypes.c_char_p), ("fdloc", ctypes.c_char_p),
                ("gridout", ctypes.c_char_p), ("media", ctypes.c_char_p), ("out", ctypes.c_char_p),
                ("spacing", ctypes.c_float), ("nt", ctypes.c_int), ("dt", ctypes.c_float),
                ("sgt_tinc", ctypes.c_int), ("component", ctypes.c_char), ("moment", ctypes.c_float),
                ("flo", ctypes.c_float), ("mu_corr", ctypes.c_int)]


class AwpError(RuntimeError):
    pass


class SgtScan(ctypes.Structure):
    """awp_sgt_scan_result (include/awpsgt.h)."""
    _fields_ = [("status", ctypes.c_int), ("offset", ctypes.c_longlong), ("nvalues", ctypes.c_longlong)]


SGT_CMP_BINS = 11
SGT_CMP_CUTOFFS = [0.0001, 0.001, 0.01, 0.1, 0.3, 1, 3, 10, 100, 1000]


class SgtCmp(ctypes.Structure):
    """awp_sgt_cmp (include/awpsgt.h)."""
    _fields_ = [("nvalues", ctypes.c_longlong), ("avg_diff", ctypes.c_double), ("max_diff", ctypes.c_float),
                ("max_diff_index", ctypes.c_longlong), ("max_abs_diff", ctypes.c_float),
                ("pct_count", ctypes.c_longlong), ("avg_pct", ctypes.c_double), ("avg_abs_pct", ctypes.c_double),
                ("max_pct", ctypes.c_float), ("max_pct_index", ctypes.c_longlong),
                ("bins", ctypes.c_longlong * SGT_CMP_BINS), ("max_series_norm", ctypes.c_double),
                ("series_checked", ctypes.c_longlong), ("series_below_floor", ctypes.c_longlong),
                ("worst_rel_l2", ctypes.c_double), ("worst_series", ctypes.c_longlong),
                ("worst_abs_below_floor", ctypes.c_double)]

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "bins"}
        d["bins"] = list(self.bins)
        return d


_lib = None

# every symbol include/awpsgt.h declares
EXPORTS = [
    "awp_in3d_read", "awp_in3d_validate", "awp_source_read", "awp_cordfile_count",
    "awp_cordfile_read", "awp_sgt_select", "awp_sgt_file_geometry", "awp_create",
    "awp_destroy", "awp_last_error", "awp_set_media", "awp_set_media_device",
    "awp_set_media_rwg", "awp_set_source", "awp_set_sgt_points", "awp_init", "awp_step", "awp_synchronize",
    "awp_progress", "awp_get_sgt", "awp_write_sgt_file", "awp_write_sgt_file_fp16",
    "awp_sgt16_read", "awp_sgt16_read_points", "awp_get_field", "awp_get_field_noalloc", "awp_set_timing", "awp_get_timing", "awp_device_bytes",
    "awp_write_sgt_file_rwg", "awp_device_memory", "awp_estimate_bytes", "awp_batch_create", "awp_batch_tune", "awp_batch_step",
    "awp_batch_synchronize", "awp_batch_time", "awp_batch_get_launch", "awp_batch_get_scheme",
    "awp_batch_last_error", "awp_batch_destroy", "awp_write_sgt_head", "awp_checkpoint_save",
    "awp_checkpoint_load", "awp_tune", "awp_set_scheme", "awp_get_scheme",
    "awp_get_trace", "awp_sgt_scan", "awp_md5_file", "awp_md5_bytes", "awp_md5_check", "awp_sgt_compare",
]


def lib() -> ctypes.CDLL:
    """Load libawpsgt.so; raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise AwpError("libawpsgt.so not built (%s); run awpsgt.build.build()" % LIB_PATH)
    # torch ships its own HIP runtime under the same soname as the system's
    # (libamdhip64.so.7); the first one loaded serves the whole process.  A
    # process that uses both torch (device memory for synthetic inputs) and
    # this library must load torch's first: torch fails to find the device
    # on a runtime of another release, while the library runs on either.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    vp, ip, cp, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t
    P = ctypes.POINTER(In3d)
    sig = {
        "awp_in3d_read": ([cp, P, cp, sz], ip),
        "awp_in3d_validate": ([P, cp, sz], ip),
        "awp_source_read": ([cp, ip, ip, vp, vp, cp, sz], ip),
        "awp_cordfile_count": ([cp, ctypes.POINTER(ctypes.c_int), cp, sz], ip),
        "awp_cordfile_read": ([cp, ip, vp, cp, sz], ip),
        "awp_sgt_select": ([P, ip, vp, vp, vp, cp, sz], ip),
        "awp_sgt_file_geometry": ([P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)], ip),
        "awp_create": ([ctypes.POINTER(vp), P, ip, ip, cp, sz], ip),
        "awp_destroy": ([vp], ip),
        "awp_last_error": ([vp], ctypes.c_char_p),
        "awp_set_media": ([vp, vp], ip),
        "awp_set_media_device": ([vp, vp], ip),
        "awp_set_media_rwg": ([vp, vp, vp, vp], ip),
        "awp_set_source": ([vp, ip, vp, vp], ip),
        "awp_set_sgt_points": ([vp, ip, vp, ctypes.POINTER(ctypes.c_int)], ip),
        "awp_init": ([vp], ip),
        "awp_step": ([vp, ip], ip),
        "awp_synchronize": ([vp], ip),
        "awp_progress": ([vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)], ip),
        "awp_get_sgt": ([vp, vp, sz], ip),
        "awp_write_sgt_file": ([vp, cp], ip),
        "awp_write_sgt_file_fp16": ([vp, cp], ip),
        "awp_write_sgt_file_rwg": ([vp, cp, ip], ip),
        "awp_sgt16_read": ([cp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), vp, sz, cp, sz], ip),
        "awp_sgt16_read_points": ([cp, ip, ip, ip, vp, sz, cp, sz], ip),
        "awp_get_field": ([vp, ip, ip, vp], ip),
        "awp_get_field_noalloc": ([vp, ip, ip, vp], ip),
        "awp_set_timing": ([vp, ip], ip),
        "awp_get_timing": ([vp, vp], ip),
        "awp_device_bytes": ([vp, ctypes.POINTER(ctypes.c_size_t)], ip),
        "awp_device_memory": ([ip, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)], ip),
        "awp_estimate_bytes": ([P, ip, ip, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)], ip),
        "awp_batch_create": ([ctypes.POINTER(vp), vp, ip, cp, sz], ip),
        "awp_write_sgt_head": ([ctypes.POINTER(HeadParams), cp, sz], ip),
        "awp_checkpoint_save": ([vp, cp], ip),
        "awp_checkpoint_load": ([vp, cp], ip),
        "awp_tune": ([vp, ip], ip),
        "awp_set_scheme": ([vp, ip], ip),
        "awp_get_scheme": ([vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                            ctypes.POINTER(ctypes.c_int)], ip),
        "awp_batch_tune": ([vp, ip], ip),
        "awp_batch_step": ([vp, ip], ip),
        "awp_batch_synchronize": ([vp], ip),
        "awp_batch_time": ([vp, ctypes.POINTER(ctypes.c_double)], ip),
        "awp_batch_get_launch": ([vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                  ctypes.POINTER(ctypes.c_int)], ip),
        "awp_batch_get_scheme": ([vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)], ip),
        "awp_batch_last_error": ([vp], ctypes.c_char_p),
        "awp_batch_destroy": ([vp], ip),
        "awp_get_trace": ([vp, vp, sz, vp], ip),
        "awp_sgt_scan": ([cp, ctypes.c_longlong, ctypes.POINTER(SgtScan), cp, sz], ip),
        "awp_md5_file": ([cp, cp, cp, sz], ip),
        "awp_md5_bytes": ([vp, sz, cp], ip),
        "awp_md5_check": ([cp, cp, cp, sz], ip),
        "awp_sgt_compare": ([cp, cp, ctypes.c_longlong, ip, ctypes.c_double, ctypes.POINTER(SgtCmp), cp, sz], ip),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _check(rc: int, msg: str) -> int:
    if rc < 0:
        raise AwpError(msg)
    return rc


def read_in3d(path: str) -> In3d:
    p = In3d()
    err = ctypes.create_string_buffer(512)
    _check(lib().awp_in3d_read(path.encode(), ctypes.byref(p), err, 512), err.value.decode())
    return p


def validate_in3d(p: In3d) -> None:
    err = ctypes.create_string_buffer(512)
    _check(lib().awp_in3d_validate(ctypes.byref(p), err, 512), err.value.decode())


def read_source(path: str, nsrc: int, nst: int) -> Tuple[np.ndarray, np.ndarray]:
    xyz = np.zeros((nsrc, 3), dtype=np.int32)
    stf = np.zeros((nsrc, nst, 6), dtype=np.float32)
    err = ctypes.create_string_buffer(512)
    _check(lib().awp_source_read(path.encode(), nsrc, nst, xyz.ctypes.data, stf.ctypes.data, err, 512),
           err.value.decode())
    return xyz, stf


def read_cordfile(path: str) -> np.ndarray:
    n = ctypes.c_int(0)
    err = ctypes.create_string_buffer(512)
    _check(lib().awp_cordfile_count(path.encode(), ctypes.byref(n), err, 512), err.value.decode())
    pts = np.zeros((n.value, 3), dtype=np.int32)
    _check(lib().awp_cordfile_read(path.encode(), n.value, pts.ctypes.data, err, 512), err.value.decode())
    return pts


def sgt_select(p: In3d, pts: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    pts = np.ascontiguousarray(pts, dtype=np.int32).reshape(-1, 3)
    keep = np.zeros(len(pts), dtype=np.int32)
    ijk = np.zeros((len(pts), 3), dtype=np.int32)
    err = ctypes.create_string_buffer(512)
    _check(lib().awp_sgt_select(ctypes.byref(p), len(pts), pts.ctypes.data, keep.ctypes.data,
                                ijk.ctypes.data, err, 512), err.value.decode())
    return keep, ijk


def sgt_file_geometry(p: In3d) -> Tuple[int, int]:
    a, b = ctypes.c_int(0), ctypes.c_int(0)
    _check(lib().awp_sgt_file_geometry(ctypes.byref(p), ctypes.byref(a), ctypes.byref(b)), "bad geometry")
    return a.value, b.value


def read_sgt16(path: str) -> np.ndarray:
    """Decode an FP16 SGT file (awp_write_sgt_file_fp16) to float32
    [npts][6][nsamp]."""
    n, t = ctypes.c_int(0), ctypes.c_int(0)
    err = ctypes.create_string_buffer(512)
    _check(lib().awp_sgt16_read(path.encode(), ctypes.byref(n), ctypes.byref(t), None, 0, err, 512),
           err.value.decode())
    out = np.empty((n.value, 6, t.value), dtype=np.float32)
    _check(lib().awp_sgt16_read(path.encode(), None, None, out.ctypes.data, out.size, err, 512),
           err.value.decode())
    return out


def read_sgt16_points(path: str, first: int, count: int, nt_out: int, layout: int = 0) -> np.ndarray:
    """Points [first, first+count) of an FP16 SGT file as float32
    [count][6][nt_out]; layout 0 AWP raw records, 1 RWG (reformat_awp),
    2 RWG for the Z source."""
    out = np.empty((count, 6, nt_out), dtype=np.float32)
    err = ctypes.create_string_buffer(512)
    _check(lib().awp_sgt16_read_points(path.encode(), int(first), int(count), int(layout), out.ctypes.data,
                                       out.size, err, 512), err.value.decode())
    return out


def sgt_scan(path: str, zero_run: int = 24000) -> Tuple[int, int, int]:
    """NaN / zero-run scan of a raw SGT file (SgtTest check_for_nans):
    (status 0 clean / 2 NaN / 3 zero run, byte offset or -1, floats read)."""
    r = SgtScan()
    err = ctypes.create_string_buffer(512)
    _check(lib().awp_sgt_scan(path.encode(), int(zero_run), ctypes.byref(r), err, 512), err.value.decode())
    return r.status, r.offset, r.nvalues


def md5_file(path: str) -> str:
    hexd = ctypes.create_string_buffer(33)
    err = ctypes.create_string_buffer(512)
    _check(lib().awp_md5_file(path.encode(), hexd, err, 512), err.value.decode())
    return hexd.value.decode()


def md5_bytes(data: bytes) -> str:
    hexd = ctypes.create_string_buffer(33)
    buf = ctypes.create_string_buffer(bytes(data), len(data))
    _check(lib().awp_md5_bytes(buf, len(data), hexd), "awp_md5_bytes")
    return hexd.value.decode()


def md5_check(path: str, md5path: str) -> bool:
    """True when the file's MD5 equals the one in md5path (CheckSgt)."""
    err = ctypes.create_string_buffer(512)
    return _check(lib().awp_md5_check(path.encode(), md5path.encode(), err, 512), err.value.decode()) == 0


def sgt_compare(ref_path: str, test_path: str, npts: int, nt: int, floor: float = 1e-6) -> dict:
    """compare_sgts statistics and the worst per-series relative L2 error of
    two raw SGT files of npts x 6 x nt float32 (see awp_sgt_compare)."""
    c = SgtCmp()
    err = ctypes.create_string_buffer(512)
    _check(lib().awp_sgt_compare(ref_path.encode(), test_path.encode(), int(npts), int(nt), float(floor),
                                 ctypes.byref(c), err, 512), err.value.decode())
    return c.as_dict()


def read_media(path: str, p: In3d) -> np.ndarray:
    n = p.nx * p.ny * p.nz * p.nvar
    return np.fromfile(path, dtype=np.float32, count=n).reshape(p.nz, p.ny, p.nx, p.nvar)


class Run:
    """One SGT component run on one GPU (an ``awp_ctx``)."""

    def __init__(self, p: In3d, device: int = 0, storage: int = AWP_STORAGE_FP32):
        self.p = p
        self.h = ctypes.c_void_p(0)
        err = ctypes.create_string_buffer(512)
        _check(lib().awp_create(ctypes.byref(self.h), ctypes.byref(p), device, storage, err, 512),
               "awp_create: " + err.value.decode())
        self.nsamp_max = p.nt // p.ntiskp_sgt
        self.npts = 0

    def _err(self) -> str:
        return lib().awp_last_error(self.h).decode()

    def _do(self, rc: int, what: str) -> int:
        if rc < 0:
            raise AwpError("%s: %s" % (what, self._err()))
        return rc

    def set_media(self, media: np.ndarray) -> None:
        m = np.ascontiguousarray(media, dtype=np.float32)
        if m.size != self.p.nx * self.p.ny * self.p.nz * self.p.nvar:
            raise ValueError("media has %d values, IN3D needs %d" % (m.size, self.p.nx * self.p.ny * self.p.nz * self.p.nvar))
        self._do(lib().awp_set_media(self.h, m.ctypes.data), "awp_set_media")

    def set_media_rwg(self, vp_km: np.ndarray, vs_km: np.ndarray, rho_g: np.ndarray) -> None:
        """Media from RWG mesh arrays [ny_rwg][nz][nx_rwg] (= AWP [nx][nz][ny])."""
        n = self.p.nx * self.p.ny * self.p.nz
        arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in (vp_km, vs_km, rho_g)]
        if any(a.size != n for a in arrs):
            raise ValueError("RWG arrays must hold %d values each" % n)
        self._do(lib().awp_set_media_rwg(self.h, *[a.ctypes.data for a in arrs]), "awp_set_media_rwg")

    def set_media_device(self, ptr: int) -> None:
        self._do(lib().awp_set_media_device(self.h, ctypes.c_void_p(ptr)), "awp_set_media_device")

    def set_source(self, xyz: np.ndarray, stf: np.ndarray) -> None:
        x = np.ascontiguousarray(xyz, dtype=np.int32).reshape(-1, 3)
        s = np.ascontiguousarray(stf, dtype=np.float32)
        if s.size != len(x) * self.p.nst * 6:
            raise ValueError("stf must be [nsrc][nst][6]")
        self._do(lib().awp_set_source(self.h, len(x), x.ctypes.data, s.ctypes.data), "awp_set_source")

    def set_sgt_points(self, pts: np.ndarray) -> int:
        x = np.ascontiguousarray(pts, dtype=np.int32).reshape(-1, 3)
        n = ctypes.c_int(0)
        self._do(lib().awp_set_sgt_points(self.h, len(x), x.ctypes.data, ctypes.byref(n)), "awp_set_sgt_points")
        self.npts = len(x)
        return n.value

    def init(self) -> None:
        self._do(lib().awp_init(self.h), "awp_init")

    def step(self, n: int) -> None:
        self._do(lib().awp_step(self.h, int(n)), "awp_step")

    def synchronize(self) -> None:
        self._do(lib().awp_synchronize(self.h), "awp_synchronize")

    def progress(self) -> Tuple[int, int]:
        a, b = ctypes.c_int(0), ctypes.c_int(0)
        self._do(lib().awp_progress(self.h, ctypes.byref(a), ctypes.byref(b)), "awp_progress")
        return a.value, b.value

    def sgt(self) -> np.ndarray:
        out = np.empty((self.npts, 6, self.nsamp_max), dtype=np.float32)
        self._do(lib().awp_get_sgt(self.h, out.ctypes.data, out.size), "awp_get_sgt")
        return out

    def trace(self) -> Tuple[np.ndarray, Tuple[int, int, int]]:
        """The per-step CHKJ trace: ([steps][3] u1 v1 w1, traced node i j k)."""
        ijk = np.zeros(3, dtype=np.int32)
        n = self.progress()[0]
        out = np.empty((n, 3), dtype=np.float32)
        self._do(lib().awp_get_trace(self.h, out.ctypes.data, out.size, ijk.ctypes.data), "awp_get_trace")
        return out, (int(ijk[0]), int(ijk[1]), int(ijk[2]))

    def write_sgt_file(self, path: str) -> None:
        self._do(lib().awp_write_sgt_file(self.h, path.encode()), "awp_write_sgt_file")

    def write_sgt_file_rwg(self, path: str, zsource: bool = False) -> None:
        """The raw file as SgtHead's reformat_awp would rewrite it (RWG order)."""
        self._do(lib().awp_write_sgt_file_rwg(self.h, path.encode(), 1 if zsource else 0),
                 "awp_write_sgt_file_rwg")

    def write_sgt_file_fp16(self, path: str) -> None:
        self._do(lib().awp_write_sgt_file_fp16(self.h, path.encode()), "awp_write_sgt_file_fp16")

    def checkpoint_save(self, path: str) -> None:
        """Write the evolving state (wavefield, memory variables, SGT samples,
        step count) to `path`."""
        self._do(lib().awp_checkpoint_save(self.h, path.encode()), "awp_checkpoint_save")

    def checkpoint_load(self, path: str) -> None:
        """Restore a checkpoint into this initialised run of the same set-up."""
        self._do(lib().awp_checkpoint_load(self.h, path.encode()), "awp_checkpoint_load")

    def tune(self, reps: int = 3) -> Tuple[bool, int, int]:
        """Time the candidate schemes and launch shapes, keep the fastest and
        start the run again (before its first step); returns scheme()."""
        self._do(lib().awp_tune(self.h, reps), "awp_tune")
        return self.scheme()

    def set_scheme(self, fused: bool) -> None:
        """Switch between the fused and the two-stage step (between steps)."""
        self._do(lib().awp_set_scheme(self.h, 1 if fused else 0), "awp_set_scheme")

    def scheme(self) -> Tuple[bool, int, int]:
        """(fused step in use, tile rows, z chunks)."""
        f, t, n = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._do(lib().awp_get_scheme(self.h, ctypes.byref(f), ctypes.byref(t), ctypes.byref(n)), "awp_get_scheme")
        return bool(f.value), t.value, n.value

    def field(self, name: str, kext: int = 0) -> np.ndarray:
        out = np.empty((self.p.nz + kext, self.p.ny, self.p.nx), dtype=np.float32)
        self._do(lib().awp_get_field(self.h, FIELDS.index(name), kext, out.ctypes.data), "awp_get_field")
        return out

    def field_noalloc(self, name: str, kext: int = 0) -> np.ndarray:
        """As field(), through awp_get_field_noalloc (no device allocation:
        the library's fixed device scratch stages the copy)."""
        out = np.empty((self.p.nz + kext, self.p.ny, self.p.nx), dtype=np.float32)
        self._do(lib().awp_get_field_noalloc(self.h, FIELDS.index(name), kext, out.ctypes.data),
                 "awp_get_field_noalloc")
        return out

    def set_timing(self, on: bool) -> None:
        self._do(lib().awp_set_timing(self.h, 1 if on else 0), "awp_set_timing")

    def timing(self) -> Tuple[float, float, float]:
        a = np.zeros(3, dtype=np.float64)
        self._do(lib().awp_get_timing(self.h, a.ctypes.data), "awp_get_timing")
        return float(a[0]), float(a[1]), float(a[2])

    def device_bytes(self) -> int:
        b = ctypes.c_size_t(0)
        self._do(lib().awp_device_bytes(self.h, ctypes.byref(b)), "awp_device_bytes")
        return b.value

    def close(self) -> None:
        if self.h and self.h.value:
            lib().awp_destroy(self.h)
            self.h = ctypes.c_void_p(0)

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def write_sgt_head(modelbox: str, coordfile: str, fdloc: str, gridout: str, spacing: float, nt: int, dt: float,
                   sgt_tinc: int, component: str, moment: float, flo: float, media: str, out: str,
                   mu_corr: bool = False) -> None:
    """SgtHead write_head's header file for an AWP SGT run (the same
    arguments, SgtHead/src/write_head.c:13-41)."""
    hp = HeadParams(modelbox.encode(), coordfile.encode(), fdloc.encode(), gridout.encode(), media.encode(),
                    out.encode(), float(spacing), int(nt), float(dt), int(sgt_tinc), component.encode()[:1],
                    float(moment), float(flo), 1 if mu_corr else 0)
    err = ctypes.create_string_buffer(512)
    _check(lib().awp_write_sgt_head(ctypes.byref(hp), err, 512), err.value.decode())


def estimate_bytes(p: In3d, storage: int = AWP_STORAGE_FP32, npts: int = 0) -> Tuple[int, int]:
    """(steady, peak) device bytes a run of ``p`` with ``npts`` SGT points
    needs (``awp_estimate_bytes``)."""
    a, b = ctypes.c_size_t(0), ctypes.c_size_t(0)
    _check(lib().awp_estimate_bytes(ctypes.byref(p), storage, int(npts), ctypes.byref(a), ctypes.byref(b)),
           "awp_estimate_bytes: bad arguments")
    return a.value, b.value


class Batch:
    """Initialised runs advanced together (``awp_batch``): one persistent
    launch per stage of a time step over every run's tiles (two-stage
    scheme, or the fused one when every run holds two wavefield buffers
    and steps fused), replayed from a HIP graph.  The runs stay usable
    (``sgt()``, ``write_sgt_file``) and must outlive the batch."""

    def __init__(self, runs: Sequence["Run"]):
        self.runs = list(runs)
        arr = (ctypes.c_void_p * len(self.runs))(*[r.h.value for r in self.runs])
        self.h = ctypes.c_void_p(0)
        err = ctypes.create_string_buffer(512)
        _check(lib().awp_batch_create(ctypes.byref(self.h), ctypes.cast(arr, ctypes.c_void_p), len(self.runs),
                                      err, 512), "awp_batch_create: " + err.value.decode())

    def _do(self, rc: int, what: str) -> int:
        if rc < 0:
            raise AwpError("%s: %s" % (what, lib().awp_batch_last_error(self.h).decode()))
        return rc

    def tune(self, reps: int = 2) -> Tuple[int, int, int]:
        """Pick the fastest kernel variant; re-initialises the runs."""
        self._do(lib().awp_batch_tune(self.h, int(reps)), "awp_batch_tune")
        return self.launch()

    def launch(self) -> Tuple[int, int, int]:
        a, b, c = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
        self._do(lib().awp_batch_get_launch(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)),
                 "awp_batch_get_launch")
        return a.value, b.value, c.value

    def scheme(self) -> Tuple[bool, bool]:
        """(fused one-sweep scheme?, persistent stage grids?)"""
        f, p = ctypes.c_int(0), ctypes.c_int(0)
        self._do(lib().awp_batch_get_scheme(self.h, ctypes.byref(f), ctypes.byref(p)), "awp_batch_get_scheme")
        return bool(f.value), bool(p.value)

    def fused(self) -> bool:
        """True when the batch uses the fused one-sweep scheme."""
        return self.scheme()[0]

    def step(self, n: int) -> None:
        self._do(lib().awp_batch_step(self.h, int(n)), "awp_batch_step")

    def time_ms(self) -> float:
        t = ctypes.c_double(0)
        self._do(lib().awp_batch_time(self.h, ctypes.byref(t)), "awp_batch_time")
        return t.value

    def close(self) -> None:
        if self.h and self.h.value:
            lib().awp_batch_destroy(self.h)
            self.h = ctypes.c_void_p(0)

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_dir(dirpath: str, device: int = 0, storage: int = AWP_STORAGE_FP32,
            nsteps: Optional[int] = None, tune: bool = True) -> Tuple[In3d, Run]:
    """Read a run directory's IN3D and inputs (paths relative to it), run it
    (after awp_tune unless tune=False, as the C host does) and return
    (params, run).  The SGT file is not written."""
    p = read_in3d(os.path.join(dirpath, "IN3D"))
    validate_in3d(p)
    media = read_media(os.path.join(dirpath, p.s("invel")), p)
    xyz, stf = read_source(os.path.join(dirpath, p.s("insrc")), p.nsrc, p.nst)
    pts = read_cordfile(os.path.join(dirpath, p.s("insgt")))
    r = Run(p, device, storage)
    r.set_sgt_points(pts)
    r.set_media(media)
    r.set_source(xyz, stf)
    r.init()
    if tune:
        r.tune()
    r.step(p.nt if nsteps is None else nsteps)
    r.synchronize()
    return p, r
