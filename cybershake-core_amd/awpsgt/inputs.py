"""Synthetic AWP-ODC-SGT inputs: IN3D, media, source and cordfile writers.

The on-disk formats are the reference solver's own, so the same files drive the
reference CPU solver and this package:

* IN3D — 72 positional list-directed records, one per line, each value followed
  by an ignored label, read in the order of AWP-ODC-SGT/src/io.f:327-418.  The
  label text follows the producer AWP-ODC-SGT/utils/build_IN3D.py:170-243.
* media (INVEL) — raw float32 ``[nz][ny][nx][nvar]`` (nvar 3: vp, vs, rho in m/s
  and kg/m^3; 5: + qp, qs), file plane z=0 is the free surface
  (AWP-ODC-SGT/src/structure.f:130-170).
* source (INSRC, IFAULT=0) — text: ``x y z`` (1-based, z=1 at the surface),
  then NST lines of ``axx, ayy, azz, axz, ayz, axy`` (source.f:170-198; sample
  AWP-ODC-SGT/utils/data/fx_src).
* cordfile (INSGT) — optional ``#`` lines, a count line, then ``x y z`` per
  point (io.f:419-442), sorted by (y, x, z) as SgtHead/gen_awp_cordfile.py:43
  writes it.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field, asdict
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

# Positional order of the 72 IN3D records (io.f:327-418).  Each entry is
# (key, kind) with kind in {"i", "f", "s"}.
IN3D_FIELDS: List[Tuple[str, str]] = [
    ("IGREEN", "i"), ("TMAX", "f"), ("DH", "f"), ("DT", "f"), ("NPC", "i"),
    ("ND", "i"), ("ARBC", "f"), ("PHT", "f"), ("NSRC", "i"), ("NST", "i"),
    ("NX", "i"), ("NY", "i"), ("NZ", "i"), ("NPX", "i"), ("NPY", "i"), ("NPZ", "i"),
    ("IFAULT", "i"), ("CHECKPOINT", "i"), ("ISFCVLM", "i"), ("IMD5", "i"),
    ("IVELOCITY", "i"), ("MEDIARESTART", "i"), ("NVAR", "i"), ("IOST", "i"),
    ("PARTDEG", "i"), ("IO_OPT", "i"), ("PERF_MEAS", "i"), ("IDYNA", "i"),
    ("SOCALQ", "i"), ("NVE", "i"), ("MU_S", "f"), ("MU_D", "f"),
    ("FL", "f"), ("FH", "f"), ("FP", "f"),
    ("READ_STEP", "i"), ("WRITE_STEP", "i"), ("WRITE_STEP2", "i"),
    ("NBGX", "i"), ("NEDX", "i"), ("NSKPX", "i"), ("NBGY", "i"), ("NEDY", "i"),
    ("NSKPY", "i"), ("NBGZ", "i"), ("NEDZ", "i"), ("NSKPZ", "i"), ("NTISKP", "i"),
    ("NBGX2", "i"), ("NEDX2", "i"), ("NSKPX2", "i"), ("NBGY2", "i"), ("NEDY2", "i"),
    ("NSKPY2", "i"), ("NBGZ2", "i"), ("NEDZ2", "i"), ("NSKPZ2", "i"),
    ("NTISKP2", "i"), ("NTISKP_SGT", "i"),
    ("CHKP", "s"), ("CHKJ", "s"), ("INSRC", "s"), ("INVEL", "s"), ("INSGT", "s"),
    ("SXRGO", "s"), ("SYRGO", "s"), ("SZRGO", "s"), ("SXRGO2", "s"), ("SYRGO2", "s"),
    ("SZRGO2", "s"), ("SGTGRO", "s"), ("SGSN", "s"),
]
assert len(IN3D_FIELDS) == 72


def default_in3d() -> Dict[str, object]:
    """SGT-mode defaults, taken from AWP-ODC-SGT/utils/data/IN3D.ref."""
    return {
        "IGREEN": 1, "TMAX": 1.0, "DH": 200.0, "DT": 0.01, "NPC": 0, "ND": 20,
        "ARBC": 0.92, "PHT": 0.1, "NSRC": 1, "NST": 100, "NX": 64, "NY": 64,
        "NZ": 32, "NPX": 1, "NPY": 1, "NPZ": 1, "IFAULT": 0, "CHECKPOINT": 0,
        "ISFCVLM": 0, "IMD5": 0, "IVELOCITY": 1, "MEDIARESTART": 1, "NVAR": 3,
        "IOST": 1, "PARTDEG": 1, "IO_OPT": 1, "PERF_MEAS": 1, "IDYNA": 0,
        "SOCALQ": 1, "NVE": 1, "MU_S": 0.677, "MU_D": 0.525, "FL": 0.01,
        "FH": 25.0, "FP": 0.5, "READ_STEP": 100, "WRITE_STEP": 10,
        "WRITE_STEP2": 10,
        "NBGX": 1, "NEDX": 1, "NSKPX": 1, "NBGY": 1, "NEDY": 1, "NSKPY": 1,
        "NBGZ": 1, "NEDZ": 1, "NSKPZ": 1, "NTISKP": 100000000,
        "NBGX2": 1, "NEDX2": 1, "NSKPX2": 1, "NBGY2": 1, "NEDY2": 1, "NSKPY2": 1,
        "NBGZ2": 1, "NEDZ2": 1, "NSKPZ2": 1, "NTISKP2": 100000000,
        "NTISKP_SGT": 10,
        "CHKP": "CHKP", "CHKJ": "CHKJ", "INSRC": "source", "INVEL": "media",
        "INSGT": "cordfile", "SXRGO": "SX", "SYRGO": "SY", "SZRGO": "SZ",
        "SXRGO2": "SX2", "SYRGO2": "SY2", "SZRGO2": "SZ2", "SGTGRO": "sgt",
        "SGSN": "SGSN",
    }


def _fmt_value(v: object, kind: str) -> str:
    if kind == "s":
        return "'%s'" % v
    if kind == "i":
        return "%9d" % int(v)
    return "%9s" % repr(float(v))


def write_in3d(path: str, params: Dict[str, object]) -> None:
    """Write an IN3D file: one ``value LABEL`` record per line, blank lines
    between groups as the reference's build_IN3D.py does."""
    full = default_in3d()
    for k, v in params.items():
        if k.upper() not in full:
            raise KeyError("unknown IN3D key %r" % k)
        full[k.upper()] = v
    lines = []
    for idx, (key, kind) in enumerate(IN3D_FIELDS):
        lines.append("%s %s" % (_fmt_value(full[key], kind), key))
        if key in ("TMAX", "DT", "NPC", "PHT", "NST", "NZ", "NPZ", "SOCALQ", "NVE",
                   "MU_D", "FP", "WRITE_STEP2", "NSKPZ", "NSKPZ2", "NTISKP2"):
            lines.append("")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


# ----------------------------------------------------------------------------
# media
# ----------------------------------------------------------------------------

def gardner_rho(vp: np.ndarray) -> np.ndarray:
    """Gardner density, rho = 1741 * (Vp in km/s)^0.25 kg/m^3."""
    return 1741.0 * np.power(np.asarray(vp, dtype=np.float64) / 1000.0, 0.25)


def layered_media(nx: int, ny: int, nz: int, vs_layers: Sequence[float],
                  vpvs: float = 1.73) -> np.ndarray:
    """Equal-thickness 1D layered medium, returned as float32 [nz][ny][nx][3]
    (vp, vs, rho); file plane 0 is the surface, the first layer is on top."""
    nl = len(vs_layers)
    vs_z = np.empty(nz, dtype=np.float64)
    for z in range(nz):
        vs_z[z] = vs_layers[min(nl - 1, (z * nl) // nz)]
    vp_z = vpvs * vs_z
    rho_z = gardner_rho(vp_z)
    out = np.empty((nz, ny, nx, 3), dtype=np.float32)
    out[..., 0] = vp_z[:, None, None]
    out[..., 1] = vs_z[:, None, None]
    out[..., 2] = rho_z[:, None, None]
    return out


def random_media(nx: int, ny: int, nz: int, seed: int, vs_top: float = 500.0,
                 vs_bot: float = 4500.0, pert: float = 0.05, vpvs: float = 1.73,
                 clip: float = 3.0) -> np.ndarray:
    """Depth-gradient Vs times (1 + pert*xi) with xi a smooth unit-variance
    random field clipped to +-clip; float32 [nz][ny][nx][3].  Small-grid
    companion of the device-side generator used for the large workloads."""
    rng = np.random.default_rng(seed)
    xi = rng.standard_normal((nz, ny, nx))
    # light smoothing so the field is correlated over a few cells
    for ax in range(3):
        xi = (np.roll(xi, 1, ax) + 2.0 * xi + np.roll(xi, -1, ax)) / 4.0
    xi /= max(xi.std(), 1e-12)
    xi = np.clip(xi, -clip, clip)
    zz = np.arange(nz, dtype=np.float64)
    vs0 = vs_top + (vs_bot - vs_top) * zz / max(nz - 1, 1)
    vs = vs0[:, None, None] * (1.0 + pert * xi)
    vp = vpvs * vs
    out = np.empty((nz, ny, nx, 3), dtype=np.float32)
    out[..., 0] = vp
    out[..., 1] = vs
    out[..., 2] = gardner_rho(vp)
    return out


def write_media(path: str, media: np.ndarray) -> None:
    media = np.ascontiguousarray(media, dtype=np.float32)
    assert media.ndim == 4 and media.shape[3] in (3, 5, 8)
    media.tofile(path)


# ----------------------------------------------------------------------------
# source
# ----------------------------------------------------------------------------

def gaussian_stf(nst: int, dt: float, sigma: float = 0.3, t0: float = 1.2,
                 amp: float = 1e9) -> np.ndarray:
    """f(t) = amp * exp(-(t-t0)^2 / (2 sigma^2)) sampled at t = (k-1) dt."""
    t = np.arange(nst, dtype=np.float64) * dt
    return (amp * np.exp(-((t - t0) ** 2) / (2.0 * sigma * sigma))).astype(np.float32)


def write_source(path: str, xyz: Tuple[int, int, int], stf: np.ndarray,
                 column: int = 0) -> None:
    """Point source file: ``x y z`` line then NST lines of six values; the
    time function goes into ``column`` (0 axx, 1 ayy, 2 azz, 3 axz, 4 ayz,
    5 axy), all other columns are zero."""
    stf = np.asarray(stf, dtype=np.float32)
    with open(path, "w") as f:
        f.write("%d %d %d\n" % tuple(int(c) for c in xyz))
        cols = ["0.0"] * 6
        for v in stf:
            cols[column] = "%.9e" % float(v)
            f.write(", ".join(cols) + "\n")


# ----------------------------------------------------------------------------
# cordfile
# ----------------------------------------------------------------------------

def random_cordfile_points(n: int, seed: int, xr: Tuple[int, int],
                           yr: Tuple[int, int], zr: Tuple[int, int]) -> np.ndarray:
    """n unique integer points (x, y, z), inclusive ranges, sorted by (y, x, z)."""
    rng = np.random.default_rng(seed)
    nxr = xr[1] - xr[0] + 1
    nyr = yr[1] - yr[0] + 1
    nzr = zr[1] - zr[0] + 1
    total = nxr * nyr * nzr
    if n > total:
        raise ValueError("more points requested than cells in range")
    if total <= 4 * n or total < 50_000_000:
        flat = rng.choice(total, size=n, replace=False)
    else:
        flat = np.unique(rng.integers(0, total, size=int(n * 1.1) + 16))
        while flat.size < n:
            flat = np.unique(np.concatenate([flat, rng.integers(0, total, size=n)]))
        flat = rng.permutation(flat)[:n]
    x = flat % nxr + xr[0]
    y = (flat // nxr) % nyr + yr[0]
    z = flat // (nxr * nyr) + zr[0]
    pts = np.stack([x, y, z], axis=1).astype(np.int64)
    order = np.lexsort((pts[:, 2], pts[:, 0], pts[:, 1]))
    return pts[order]


def write_cordfile(path: str, pts: np.ndarray, header: Optional[str] = None) -> None:
    pts = np.asarray(pts, dtype=np.int64)
    with open(path, "w") as f:
        if header:
            for line in header.splitlines():
                f.write("# %s\n" % line)
        f.write("%d\n" % len(pts))
        np.savetxt(f, pts, fmt="%d")


# ----------------------------------------------------------------------------
# a complete run directory
# ----------------------------------------------------------------------------

@dataclass
class RunSpec:
    """One SGT component run: grid, medium, source and output points."""
    nx: int
    ny: int
    nz: int
    dh: float
    dt: float
    nt: int                       # steps; TMAX is chosen so nint(TMAX/DT)+1 == nt
    igreen: int = 1
    src: Tuple[int, int, int] = (1, 1, 1)
    sigma: float = 0.3
    t0: float = 1.2
    amp: float = 1e9
    nd: int = 20
    arbc: float = 0.92
    socalq: int = 1
    fl: float = 0.01
    fh: float = 25.0
    fp: float = 0.5
    ntiskp_sgt: int = 10
    write_step: Optional[int] = None
    npx: int = 1
    npy: int = 1
    npz: int = 1
    nvar: int = 3
    extra: Dict[str, object] = field(default_factory=dict)

    def tmax(self) -> float:
        return (self.nt - 1) * self.dt

    def nst(self) -> int:
        return self.nt

    def ws(self) -> int:
        return self.write_step if self.write_step else self.nt // self.ntiskp_sgt


def write_run(dirpath: str, spec: RunSpec, media: np.ndarray, pts: np.ndarray) -> Dict[str, str]:
    """Write IN3D, media, source and cordfile for ``spec`` into ``dirpath``."""
    os.makedirs(dirpath, exist_ok=True)
    column = {1: 0, 2: 1, 3: 2, 4: 0, 5: 1, 6: 2}[spec.igreen]
    stf = gaussian_stf(spec.nst(), spec.dt, spec.sigma, spec.t0, spec.amp)
    write_media(os.path.join(dirpath, "media"), media)
    write_source(os.path.join(dirpath, "source"), spec.src, stf, column)
    write_cordfile(os.path.join(dirpath, "cordfile"), pts)
    params = {
        "IGREEN": spec.igreen, "TMAX": float(np.float32(spec.tmax())), "DH": spec.dh,
        "DT": spec.dt, "ND": spec.nd, "ARBC": spec.arbc, "NST": spec.nst(),
        "READ_STEP": spec.nst(), "NX": spec.nx, "NY": spec.ny, "NZ": spec.nz,
        "NPX": spec.npx, "NPY": spec.npy, "NPZ": spec.npz, "NVAR": spec.nvar,
        "SOCALQ": spec.socalq, "FL": spec.fl, "FH": spec.fh, "FP": spec.fp,
        "WRITE_STEP": spec.ws(), "WRITE_STEP2": spec.ws(),
        "NTISKP_SGT": spec.ntiskp_sgt,
    }
    params.update(spec.extra)
    write_in3d(os.path.join(dirpath, "IN3D"), params)
    return {"in3d": os.path.join(dirpath, "IN3D"), "sgt": os.path.join(dirpath, "sgt")}
