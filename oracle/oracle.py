"""Python front end of the CPU restatement (oracle/awp_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product never imports it.

Also restates the reference's input readers so tests can drive the oracle from
the same files the product reads:

* ``read_in3d``   -- io.f:327-418 (72 list-directed records, one per line;
  blank lines skipped; anything after the value ignored).
* ``read_source`` -- source.f:170-198 (IFAULT=0 text source).
* ``read_cordfile`` -- io.f:419-442 (leading '#' lines skipped).
* ``read_media``  -- structure.f:130-170 (float32 [nz][ny][nx][nvar]).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Dict, List, Optional, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libawp_oracle.so")

# positional order of io.f:327-418
_IN3D_KEYS = [
    ("IGREEN", int), ("TMAX", float), ("DH", float), ("DT", float), ("NPC", int),
    ("ND", int), ("ARBC", float), ("PHT", float), ("NSRC", int), ("NST", int),
    ("NX", int), ("NY", int), ("NZ", int), ("NPX", int), ("NPY", int), ("NPZ", int),
    ("IFAULT", int), ("CHECKPOINT", int), ("ISFCVLM", int), ("IMD5", int),
    ("IVELOCITY", int), ("MEDIARESTART", int), ("NVAR", int), ("IOST", int),
    ("PARTDEG", int), ("IO_OPT", int), ("PERF_MEAS", int), ("IDYNA", int),
    ("SOCALQ", int), ("NVE", int), ("MU_S", float), ("MU_D", float),
    ("FL", float), ("FH", float), ("FP", float),
    ("READ_STEP", int), ("WRITE_STEP", int), ("WRITE_STEP2", int),
    ("NBGX", int), ("NEDX", int), ("NSKPX", int), ("NBGY", int), ("NEDY", int),
    ("NSKPY", int), ("NBGZ", int), ("NEDZ", int), ("NSKPZ", int), ("NTISKP", int),
    ("NBGX2", int), ("NEDX2", int), ("NSKPX2", int), ("NBGY2", int), ("NEDY2", int),
    ("NSKPY2", int), ("NBGZ2", int), ("NEDZ2", int), ("NSKPZ2", int),
    ("NTISKP2", int), ("NTISKP_SGT", int),
    ("CHKP", str), ("CHKJ", str), ("INSRC", str), ("INVEL", str), ("INSGT", str),
    ("SXRGO", str), ("SYRGO", str), ("SZRGO", str), ("SXRGO2", str), ("SYRGO2", str),
    ("SZRGO2", str), ("SGTGRO", str), ("SGSN", str),
]


def _first_token(line: str) -> str:
    s = line.strip()
    if s[:1] in ("'", '"'):
        q = s[0]
        end = s.find(q, 1)
        return s[1:end] if end > 0 else s[1:]
    for sep in (" ", "\t", ","):
        pos = s.find(sep)
        if pos >= 0:
            s = s[:pos]
    return s


def read_in3d(path: str) -> Dict[str, object]:
    vals: Dict[str, object] = {}
    with open(path) as f:
        lines = [ln for ln in f.read().splitlines() if ln.strip()]
    if len(lines) < len(_IN3D_KEYS):
        raise ValueError("IN3D has %d records, expected %d" % (len(lines), len(_IN3D_KEYS)))
    for (key, typ), ln in zip(_IN3D_KEYS, lines):
        tok = _first_token(ln)
        if typ is int:
            vals[key] = int(tok)
        elif typ is float:
            vals[key] = float(np.float32(float(tok.replace("d", "e").replace("D", "e"))))
        else:
            vals[key] = tok
    # nt = nint(tmax/dt) + 1 in single precision (pmcl3d.f:287)
    q = np.float32(vals["TMAX"]) / np.float32(vals["DT"])
    vals["NT"] = int(np.floor(np.float32(q) + np.float32(0.5))) + 1 if q >= 0 else int(np.ceil(q - 0.5)) + 1
    return vals


def read_source(path: str, nsrc: int, nst: int, nz: Optional[int] = None) -> Tuple[np.ndarray, np.ndarray]:
    """Returns (xyz[nsrc,3] in file coordinates, stf[nsrc,nst,6] float32)."""
    xyz = np.zeros((nsrc, 3), dtype=np.int32)
    stf = np.zeros((nsrc, nst, 6), dtype=np.float32)
    with open(path) as f:
        lines = [ln for ln in f.read().splitlines() if ln.strip()]
    pos = 0
    for s in range(nsrc):
        xyz[s] = [int(t) for t in lines[pos].replace(",", " ").split()[:3]]
        pos += 1
        for t in range(nst):
            stf[s, t] = [np.float32(float(v)) for v in lines[pos].replace(",", " ").split()[:6]]
            pos += 1
    return xyz, stf


def read_cordfile(path: str) -> np.ndarray:
    with open(path) as f:
        lines = f.read().splitlines()
    i = 0
    while lines[i].strip().startswith("#") or not lines[i].strip():
        i += 1
    n = int(lines[i].split()[0])
    pts = np.array([[int(t) for t in lines[i + 1 + q].split()[:3]] for q in range(n)], dtype=np.int32)
    return pts.reshape(n, 3)


def read_media(path: str, nx: int, ny: int, nz: int, nvar: int) -> np.ndarray:
    return np.fromfile(path, dtype=np.float32, count=nx * ny * nz * nvar).reshape(nz, ny, nx, nvar)


class _Params(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("npx", ctypes.c_int), ("npy", ctypes.c_int), ("npz", ctypes.c_int),
        ("dh", ctypes.c_double), ("dt", ctypes.c_double),
        ("nt", ctypes.c_int), ("nd", ctypes.c_int), ("arbc", ctypes.c_double),
        ("socalq", ctypes.c_int), ("fl", ctypes.c_double), ("fh", ctypes.c_double),
        ("fp", ctypes.c_double), ("igreen", ctypes.c_int), ("nst", ctypes.c_int),
        ("ntiskp_sgt", ctypes.c_int), ("write_step", ctypes.c_int), ("nvar", ctypes.c_int),
    ]


_lib = None


def build() -> None:
    subprocess.check_call(["sh", os.path.join(HERE, "build_oracle.sh")])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        for sfx in ("f32", "f64", "mix"):
            getattr(L, "awp_oracle_create_" + sfx).restype = ctypes.c_void_p
            getattr(L, "awp_oracle_create_" + sfx).argtypes = [
                ctypes.POINTER(_Params), ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
            getattr(L, "awp_oracle_destroy_" + sfx).argtypes = [ctypes.c_void_p]
            getattr(L, "awp_oracle_run_" + sfx).argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
            getattr(L, "awp_oracle_field_" + sfx).argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
            getattr(L, "awp_oracle_sgt_" + sfx).argtypes = [ctypes.c_void_p, ctypes.c_void_p]
            getattr(L, "awp_oracle_info_" + sfx).argtypes = [ctypes.c_void_p, ctypes.c_void_p]
            getattr(L, "awp_oracle_points_" + sfx).argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        _lib = L
    return _lib


FIELDS = ["u1", "v1", "w1", "xx", "yy", "zz", "xy", "xz", "yz",
          "r1", "r2", "r3", "r4", "r5", "r6", "d1", "mu", "lam", "qp", "qs"]


class Oracle:
    """One SGT run of the CPU restatement.  ``dtype`` float32 (reference
    operation order) or float64 (FP64 companion); ``mixed=True`` selects the
    restatement of the device's mixed-storage mode (FP32 arithmetic, FP16
    memory variables under per-row exponents, FP16 qp/qs, FP16 d1/mu/lam
    deviations from per-level profiles)."""

    def __init__(self, in3d: Dict[str, object], media: np.ndarray, src_xyz: np.ndarray,
                 stf: np.ndarray, pts: np.ndarray, dtype=np.float32, mixed: bool = False):
        self.dtype = np.dtype(dtype)
        if mixed and self.dtype != np.float32:
            raise ValueError("the mixed-storage restatement computes in float32")
        self.sfx = "mix" if mixed else ("f32" if self.dtype == np.float32 else "f64")
        L = lib()
        p = _Params()
        p.nx, p.ny, p.nz = int(in3d["NX"]), int(in3d["NY"]), int(in3d["NZ"])
        p.npx, p.npy, p.npz = int(in3d.get("NPX", 1)), int(in3d.get("NPY", 1)), int(in3d.get("NPZ", 1))
        p.dh, p.dt = float(in3d["DH"]), float(in3d["DT"])
        p.nt = int(in3d["NT"])
        p.nd = int(in3d["ND"])
        p.arbc = float(in3d["ARBC"])
        p.socalq = int(in3d["SOCALQ"])
        p.fl, p.fh, p.fp = float(in3d["FL"]), float(in3d["FH"]), float(in3d["FP"])
        p.igreen = int(in3d["IGREEN"])
        p.nst = int(in3d["NST"])
        p.ntiskp_sgt = int(in3d["NTISKP_SGT"])
        p.write_step = int(in3d["WRITE_STEP"])
        p.nvar = int(in3d["NVAR"])
        self.p = p
        self._media = np.ascontiguousarray(media, dtype=np.float32)
        self._src = np.ascontiguousarray(src_xyz, dtype=np.int32).reshape(-1, 3)
        self._stf = np.ascontiguousarray(stf, dtype=np.float32)
        self._pts = np.ascontiguousarray(pts, dtype=np.int32).reshape(-1, 3)
        self.h = getattr(L, "awp_oracle_create_" + self.sfx)(
            ctypes.byref(p), self._media.ctypes.data, len(self._src), self._src.ctypes.data,
            self._stf.ctypes.data, len(self._pts), self._pts.ctypes.data)
        if not self.h:
            raise ValueError("oracle rejected the inputs")
        self.nsamp_max = p.nt // p.ntiskp_sgt

    def run(self, nsteps: int, nthreads: int = 0) -> int:
        return getattr(lib(), "awp_oracle_run_" + self.sfx)(self.h, int(nsteps), int(nthreads))

    def field(self, name: str, kext: int = 0) -> np.ndarray:
        out = np.empty((self.p.nz + kext, self.p.ny, self.p.nx), dtype=self.dtype)
        rc = getattr(lib(), "awp_oracle_field_" + self.sfx)(self.h, FIELDS.index(name), kext, out.ctypes.data)
        if rc != 0:
            raise ValueError(name)
        return out

    def sgt(self) -> Tuple[np.ndarray, int]:
        out = np.empty((len(self._pts), 6, max(self.nsamp_max, 1)), dtype=self.dtype)
        n = getattr(lib(), "awp_oracle_sgt_" + self.sfx)(self.h, out.ctypes.data)
        return out, n

    def info(self) -> Dict[str, int]:
        a = np.zeros(4, dtype=np.int32)
        getattr(lib(), "awp_oracle_info_" + self.sfx)(self.h, a.ctypes.data)
        return {"step": int(a[0]), "nsamp": int(a[1]), "nsamp_max": int(a[2]), "nkept": int(a[3])}

    def points(self) -> Tuple[np.ndarray, np.ndarray]:
        n = self.info()["nkept"]
        kidx = np.zeros(max(n, 1), dtype=np.int32)
        kpos = np.zeros((max(n, 1), 3), dtype=np.int32)
        getattr(lib(), "awp_oracle_points_" + self.sfx)(self.h, kidx.ctypes.data, kpos.ctypes.data)
        return kidx[:n], kpos[:n]

    def close(self) -> None:
        if self.h:
            getattr(lib(), "awp_oracle_destroy_" + self.sfx)(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sgt_file_layout(nt: int, ntiskp: int, ws: int) -> Tuple[int, List[int]]:
    """Raw SGT file geometry of pmcl3d.f:1536-1601: returns (NT_OUT, list of
    step indices at which a WRITE_STEP buffer is flushed)."""
    nt_out = int(int(nt / ws) * ws / ntiskp)
    flushes = []
    count = 0
    for i in range(1, nt + 1):
        if i == ntiskp * (i // ntiskp):
            count += 1
            if count == ws:
                flushes.append(i)
                count = 0
    return nt_out, flushes


def sgt_file_from_samples(samples: np.ndarray, nt: int, ntiskp: int, ws: int,
                          kept: Optional[np.ndarray] = None) -> Optional[bytes]:
    """Bytes of the raw SGT file the reference writes for the given samples
    [npts][6][nsamp] (pmcl3d.f:1582-1601): only full WRITE_STEP buffers are
    written, each at offset g*24*NT_OUT + c*4*NT_OUT + (f-1)*4*WS; points not
    kept leave holes (zero bytes).  None when no buffer is ever flushed."""
    npts = samples.shape[0]
    nt_out, flushes = sgt_file_layout(nt, ntiskp, ws)
    if not flushes:
        return None
    kept_set = np.arange(npts) if kept is None else np.asarray(kept)
    size = 0
    chunks = []
    for fi, step in enumerate(flushes):
        f = int(step / ntiskp / ws)
        s0 = fi * ws
        for g in kept_set:
            for c in range(6):
                off = int(g) * 6 * 4 * nt_out + c * 4 * nt_out + (f - 1) * 4 * ws
                chunks.append((off, samples[g, c, s0:s0 + ws].astype(np.float32)))
                size = max(size, off + 4 * ws)
    buf = bytearray(size)
    for off, vals in chunks:
        buf[off:off + 4 * len(vals)] = vals.tobytes()
    return bytes(buf)
