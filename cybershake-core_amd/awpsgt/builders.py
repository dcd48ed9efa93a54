"""Input builders of a CyberShake SGT run: from the RWG-convention inputs a
site produces (velocity mesh, SGT point list, grid file, source location)
to the AWP-ODC-SGT inputs the solver reads.  Each mirrors a reference tool;
the reference's x/y swap (RWG x = AWP y) and 1-based AWP indexing are kept.

    python -m awpsgt.builders cordfile <rwg cordfile> <awp cordfile> <max depth index>
    python -m awpsgt.builders velocity <nx> <ny> <nz> <rwg prefix> <awp media out>
    python -m awpsgt.builders src <site> <fdloc file> <x|y|z> <stf file> <out>
    python -m awpsgt.builders in3d <site> <gridout> <x|y|z> <frequency> <px> <py> <pz> <mesh name> [tmax]
    python -m awpsgt.builders head <modelbox> <coordfile> <fdloc> <gridout> <spacing> <nt> <dt> <tinc>
                                   <x|y|z> <moment> <max freq> <awp media> <out> [-c]

The last is the output-side companion (SgtHead write_head: the header file
of an SGT run, through awp_write_sgt_head in the C ABI).

The velocity reformat also has a device path (awp_set_media_rwg in the C
ABI) that takes the three RWG arrays straight into a run's media set-up.
"""
from __future__ import annotations

import sys
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np


# ---------------------------------------------------------------------------
# SGT point list: SgtHead/gen_awp_cordfile.py:1-52
# ---------------------------------------------------------------------------
def rwg_to_awp_points(lines: Sequence[str]) -> Tuple[List[Tuple[int, int, int]], int, List[str]]:
    """RWG cordfile text -> (AWP points, points the input declares, duplicate
    messages).  Line 5 (index 4) holds the declared count, points follow as
    "x y z ..." (gen_awp_cordfile.py:24-35); each becomes (y+1, x+1, z) (z is
    1 at the free surface in both conventions); duplicates are dropped and
    reported; the result is sorted by (AWP y, AWP x, z) (:37-41)."""
    n_in = int(lines[4])
    seen = set()
    dups = []
    for ln in lines[5:]:
        pc = ln.split()
        if len(pc) < 3:
            # the reference would stop with an error here; a blank tail is skipped
            if pc:
                raise ValueError("cordfile line %r has fewer than 3 fields" % ln)
            continue
        p = (int(pc[1]) + 1, int(pc[0]) + 1, int(pc[2]))
        if p in seen:
            dups.append("Duplicate point entry %d %d %d" % p)
        seen.add(p)
    pts = sorted(seen, key=lambda q: (q[1], q[0], q[2]))
    return pts, n_in, dups


def awp_cordfile_text(pts: Sequence[Tuple[int, int, int]]) -> str:
    """The AWP cordfile exactly as gen_awp_cordfile.py writes it (:42-47):
    the count, then one "x y z " line (trailing space) per point."""
    out = ["%d\n" % len(pts)]
    for p in pts:
        out.append("".join("%d " % v for v in p) + "\n")
    return "".join(out)


def gen_awp_cordfile(rwg_path: str, awp_path: str) -> int:
    """File-to-file form; returns 2 when the point count changed (duplicates),
    as the reference's exit status (:49-52), else 0."""
    with open(rwg_path) as f:
        lines = f.readlines()
    pts, n_in, dups = rwg_to_awp_points(lines)
    for d in dups:
        print(d)
    with open(awp_path, "w") as f:
        f.write(awp_cordfile_text(pts))
    if n_in != len(pts):
        print("Error: input file %s has %d points, but %d points were written to output file %s."
              % (rwg_path, n_in, len(pts), awp_path))
        return 2
    return 0


# ---------------------------------------------------------------------------
# velocity mesh: SgtHead/src/reformat_velocity.c:1-95
# ---------------------------------------------------------------------------
def reformat_velocity_arrays(p: np.ndarray, s: np.ndarray, d: np.ndarray) -> np.ndarray:
    """RWG arrays [ny][nz][nx] (x fastest, then z, then y; km/s, g/cm^3) ->
    AWP media [nz][nx][ny][3] (vp, vs, rho in m/s, kg/m^3), i.e. the AWP file
    [z][y_awp][x_awp][3] with x_awp = RWG y (reformat_velocity.c:62-72); each
    value is the float nearest value * 1000 (the reference multiplies in
    double and stores float, which rounds once like a float multiply)."""
    out = np.empty(p.shape[1:2] + (p.shape[2], p.shape[0], 3), dtype=np.float32)
    for c, a in enumerate((p, s, d)):
        a32 = np.asarray(a, dtype=np.float32)
        out[..., c] = (np.transpose(a32, (1, 2, 0)).astype(np.float64) * 1000.0).astype(np.float32)
    return out


def reformat_velocity(nx: int, ny: int, nz: int, prefix: str, out_path: str, chunk_y: int = 64) -> None:
    """File-to-file form over <prefix>.p/.s/.d, a band of RWG y at a time."""
    with open(prefix + ".p", "rb") as fp, open(prefix + ".s", "rb") as fs, open(prefix + ".d", "rb") as fd, \
            open(out_path, "wb") as fo:
        # the output is [nz][nx][ny][3]: every band of y writes into every (z, x) row
        fo.truncate(nz * nx * ny * 3 * 4)
        for y0 in range(0, ny, chunk_y):
            m = min(chunk_y, ny - y0)
            cnt = nx * nz * m
            arrs = [np.fromfile(f, dtype=np.float32, count=cnt).reshape(m, nz, nx) for f in (fp, fs, fd)]
            band = reformat_velocity_arrays(*arrs)   # [nz][nx][m][3]
            for k in range(nz):
                for i in range(nx):
                    fo.seek(4 * 3 * (k * nx * ny + i * ny + y0))
                    fo.write(band[k, i].tobytes())


# ---------------------------------------------------------------------------
# source file: AWP-GPU-SGT/utils/build_src.py:6-46
# ---------------------------------------------------------------------------
def awp_source_text(fdloc_line: str, stf_lines: Sequence[str]) -> str:
    """"<y+1> <x+1> 1" (the RWG source location from the fdloc file, x/y
    swapped, 1-based, at the surface index 1) followed by the STF lines as
    they are (build_src.py:38-43)."""
    src_x, src_y = fdloc_line.split()[:2]
    return "%d %d 1\n" % (int(src_y) + 1, int(src_x) + 1) + "".join(stf_lines)


def build_src(fdloc_path: str, stf_path: str, out_path: str) -> None:
    with open(fdloc_path) as f:
        first = f.readline()
    with open(stf_path) as f:
        stf = f.readlines()
    with open(out_path, "w") as f:
        f.write(awp_source_text(first, stf))


# ---------------------------------------------------------------------------
# IN3D: AWP-GPU-SGT/utils/build_IN3D.py:75-244 with data/IN3D.ref defaults
# ---------------------------------------------------------------------------
# the defaults of AWP-GPU-SGT/utils/data/IN3D.ref (one value per key)
IN3D_REF: Dict[str, str] = {
    "TMAX": "200.0", "DH": "200.0", "DT": "0.01", "NPC": "0", "ND": "80", "ARBC": "0.92", "PHT": "0.1",
    "NSRC": "1", "NST": "20000", "IFAULT": "0", "CHECKPOINT": "0", "ISFCVLM": "0", "IMD5": "0",
    "IVELOCITY": "1", "MEDIARESTART": "2", "NVAR": "3", "IOST": "1", "PARTDEG": "1", "IO_OPT": "1",
    "PERF_MEAS": "1", "IDYNA": "0", "SOCALQ": "1", "NVE": "1", "MU_S": "0.677", "MU_D": "0.525",
    "FL": "0.01", "FH": "25.0", "FP": "1.0", "READ_STEP": "20000", "WRITE_STEP": "2000",
    "WRITE_STEP2": "2000", "NSKPX": "10", "NSKPY": "10", "NSKPZ": "10", "NTISKP": "100000",
    "NSKPX2": "10", "NSKPY2": "10", "NSKPZ2": "10", "NTISKP2": "1000000", "NTISKP_SGT": "10",
    "CHKP": "output_ckp/CHKP96PV", "CHKJ": "output_ckp/CHKJ96PV", "SXRGO": "output_sfc/XX96PS",
    "SYRGO": "output_sfc/YY96PS", "SZRGO": "output_sfc/ZZ96PS", "SXRGO2": "output_sfc/XY96PV",
    "SYRGO2": "output_sfc/YZ96PV", "SZRGO2": "output_sfc/XZ96PV", "SGSN": "output_dyn/SGSN",
}


def gridout_dims(lines: Sequence[str]) -> Tuple[int, int, int]:
    """AWP (nx, ny, nz) from an RWG gridout file: its first block is RWG x
    (= AWP y), then y, then z, each "n=<count>" followed by count + 1 lines
    (build_IN3D.py:120-127)."""
    ny = int(lines[1].split("=")[1])
    nx = int(lines[1 + ny + 2].split("=")[1])
    nz = int(lines[1 + ny + 2 + nx + 2].split("=")[1])
    return nx, ny, nz


def in3d_params(site: str, dims: Tuple[int, int, int], awp_comp: str, frequency: float,
                proc: Tuple[int, int, int], mesh_filename: str, spacing: Optional[float] = None,
                ref: Optional[Dict[str, str]] = None) -> Dict[str, object]:
    """The parameter set build_IN3D.py computes (:75-168)."""
    param: Dict[str, object] = dict(ref or IN3D_REF)
    igreen = {"x": 1, "y": 2, "z": 3}.get(awp_comp)
    if igreen is None:
        raise ValueError("component %r not recognized" % awp_comp)
    param["igreen"] = igreen
    if spacing is not None:
        param["DH"] = 1000.0 * spacing
        param["DT"] = spacing / 20.0
    else:
        param["DH"] = round(100.0 / frequency, 1)
        param["DT"] = 0.005 / frequency
    sim_time = float(param["TMAX"])
    nst = int(sim_time / param["DT"])          # rounded up to a multiple of 1000
    if nst % 1000 != 0:
        nst = 1000 * (nst // 1000 + 1)
    param["NST"] = nst
    param["TMAX"] = nst * param["DT"]
    param["FP"] = 0.5
    param["READ_STEP"] = nst
    param["WRITE_STEP"] = nst // int(param["NTISKP_SGT"])
    param["WRITE_STEP2"] = param["WRITE_STEP"]
    nx, ny, nz = dims
    for n, pr, ax in ((nx, proc[0], "X"), (ny, proc[1], "Y"), (nz, proc[2], "Z")):
        if n % pr != 0:
            raise ValueError("P%s %d must be a factor of N%s %d" % (ax, pr, ax, n))
    param.update(NX=nx, NY=ny, NZ=nz, NPX=proc[0], NPY=proc[1], NPZ=proc[2],
                 NBGX=1, NEDX=nx, NBGY=1, NEDY=ny, NBGZ=1, NEDZ=nz,
                 NBGX2=1, NEDX2=nx, NBGY2=1, NEDY2=ny, NBGZ2=1, NEDZ2=nz)
    param["INSRC"] = "comp_%s/input/%s_f%s_src" % (awp_comp, site, awp_comp)
    param["INVEL"] = "comp_%s/input/%s" % (awp_comp, mesh_filename)
    param["INSGT"] = "comp_%s/input/awp.%s.cordfile" % (awp_comp, site)
    param["SGTGRO"] = "comp_%s/output_sgt/awp-strain-%s-f%s" % (awp_comp, site, awp_comp)
    return param


def in3d_text(param: Dict[str, object], awp_comp: str) -> str:
    """IN3D in build_IN3D.py's layout (:171-242): value right-aligned in 9
    columns, the key after it, blank lines between groups, quoted paths."""
    P = param
    w = []

    def v(key, fmt="%9s"):
        w.append((fmt % (P[key],)) + " " + key + "\n")

    v("igreen"); v("TMAX"); w.append("\n")
    v("DH"); v("DT"); w.append("\n")
    v("NPC"); w.append("\n")
    v("ND"); v("ARBC"); v("PHT"); w.append("\n")
    v("NSRC"); v("NST"); w.append("\n")
    v("NX", "%9d"); v("NY", "%9d"); v("NZ", "%9d"); w.append("\n")
    v("NPX", "%9d"); v("NPY", "%9d"); v("NPZ", "%9d"); w.append("\n")
    for k in ("IFAULT", "CHECKPOINT", "ISFCVLM", "IMD5", "IVELOCITY", "MEDIARESTART", "NVAR", "IOST",
              "PARTDEG", "IO_OPT", "PERF_MEAS", "IDYNA", "SOCALQ"):
        v(k)
    w.append("\n")
    v("NVE"); w.append("\n")
    v("MU_S"); v("MU_D"); w.append("\n")
    v("FL"); v("FH"); v("FP"); w.append("\n")
    v("READ_STEP"); v("WRITE_STEP"); v("WRITE_STEP2"); w.append("\n")
    for k in ("NBGX", "NEDX", "NSKPX", "NBGY", "NEDY", "NSKPY", "NBGZ", "NEDZ", "NSKPZ"):
        v(k)
    w.append("\n")
    # NTISKP beyond the step count: no velocity output (build_IN3D.py:216)
    w.append("%9s NTISKP\n" % (2 * int(P["NST"]),))
    for k in ("NBGX2", "NEDX2", "NSKPX2", "NBGY2", "NEDY2", "NSKPY2", "NBGZ2", "NEDZ2", "NSKPZ2"):
        v(k)
    w.append("\n")
    w.append("%9s NTISKP2\n\n" % (P["NTISKP"],))
    v("NTISKP_SGT"); w.append("\n")
    w.append("'comp_%s/%s' CHKP\n" % (awp_comp, P["CHKP"]))
    w.append("'comp_%s/%s' CHKJ\n\n" % (awp_comp, P["CHKJ"]))
    w.append("'%s' INSRC\n" % (P["INSRC"],))
    w.append("'%s' INVEL\n\n" % (P["INVEL"],))
    w.append("'%s' INSGT\n\n" % (P["INSGT"],))
    for k in ("SXRGO", "SYRGO", "SZRGO"):
        w.append("'comp_%s/%s' %s\n" % (awp_comp, P[k], k))
    w.append("\n")
    for k in ("SXRGO2", "SYRGO2", "SZRGO2"):
        w.append("'comp_%s/%s' %s\n" % (awp_comp, P[k], k))
    w.append("\n")
    w.append("'%s' SGTGRO\n\n" % (P["SGTGRO"],))
    w.append("'comp_%s/%s' SGSN\n" % (awp_comp, P["SGSN"]))
    return "".join(w)


def _main(argv: Sequence[str]) -> int:
    if len(argv) < 2:
        print(__doc__)
        return 1
    cmd, a = argv[1], argv[2:]
    if cmd == "cordfile" and len(a) >= 3:
        return gen_awp_cordfile(a[0], a[1])
    if cmd == "velocity" and len(a) >= 5:
        reformat_velocity(int(a[0]), int(a[1]), int(a[2]), a[3], a[4])
        return 0
    if cmd == "src" and len(a) >= 5:
        build_src(a[1], a[3], a[4])
        return 0
    if cmd == "in3d" and len(a) >= 8:
        site, gridout, comp, freq = a[0], a[1], a[2], float(a[3])
        proc = (int(a[4]), int(a[5]), int(a[6]))
        ref = dict(IN3D_REF)
        if len(a) > 8:
            ref["TMAX"] = a[8]
        with open(gridout) as f:
            dims = gridout_dims(f.readlines())
        P = in3d_params(site, dims, comp, freq, proc, a[7], ref=ref)
        with open("IN3D.%s.%s" % (site, comp), "w") as f:
            f.write(in3d_text(P, comp))
        return 0
    if cmd == "head" and len(a) >= 13:
        import awpsgt
        awpsgt.write_sgt_head(a[0], a[1], a[2], a[3], float(a[4]), int(a[5]), float(a[6]), int(a[7]), a[8],
                              float(a[9]), float(a[10]), a[11], a[12], mu_corr=len(a) > 13 and a[13] == "-c")
        return 0
    print(__doc__)
    return 1


if __name__ == "__main__":
    sys.exit(_main(sys.argv))
