"""Seeded inputs of the SGT-header case (SgtHead write_head): a 20x16x12
RWG grid with a model box, a site location, an RWG coordfile whose points
cover the edges the header code treats specially (x = nx-1, y = 0, y = ny-1,
z = 0, z = nz-1) and an AWP-format media file ([z][x][y][3] Vp, Vs m/s, rho
kg/m^3)."""
import os

import numpy as np

NX, NY, NZ = 20, 16, 12
SPACING = 0.1
SITE = (10, 8)
ARGS = dict(spacing=SPACING, nt=2000, dt=0.005, sgt_tinc=10, moment=1.0e20, flo=1.0)


def write_inputs(d: str, seed: int = 5) -> dict:
    rng = np.random.default_rng(seed)
    paths = {k: os.path.join(d, k) for k in ("modelbox", "coordfile", "fdloc", "gridout", "media")}
    with open(paths["modelbox"], "w") as f:
        f.write("TEST\nAPPROXIMATE CENTROID:\n  clon= -118.2000 clat= 34.0500\nMODEL CORNERS:\n"
                "  mlon= -118.28600 mlat= 34.01920 mrot= -55.0 xlen= 2.0000 ylen= 1.6000\n")
    with open(paths["gridout"], "w") as f:
        for key, n in (("x", NX), ("y", NY), ("z", NZ)):
            f.write("%slen=%.4f\nn%s=%d\n" % (key, n * SPACING, key, n))
            for i in range(n):
                f.write("%5d %12.4f %12.4f\n" % (i, i * SPACING, SPACING))
    with open(paths["fdloc"], "w") as f:
        f.write("%d %d\n" % SITE)
    pts = [(NX - 1, 3, 4), (5, 0, 2), (7, NY - 1, 6), (2, 9, 0), (11, 4, NZ - 1), (0, 0, 0), (NX - 1, NY - 1, NZ - 1),
           (NX - 2, NY - 1, 3), SITE + (0,)]
    for _ in range(31):
        pts.append((int(rng.integers(0, NX)), int(rng.integers(0, NY)), int(rng.integers(0, NZ))))
    with open(paths["coordfile"], "w") as f:
        f.write("# header lines\n# of the coordfile\n%d\n" % len(pts))
        for x, y, z in pts:
            f.write("%d %d %d %d %.5f %.5f %.4f\n" % (x, y, z, x * 10 ** 12 + y * 10 ** 6 + z,
                                                      -118.0 - 0.01 * x, 34.0 + 0.01 * y, z * SPACING))
    vp = rng.uniform(1700.0, 6500.0, (NZ, NX, NY)).astype(np.float32)
    vs = (vp / rng.uniform(1.6, 2.0, (NZ, NX, NY))).astype(np.float32)
    rho = rng.uniform(1900.0, 3100.0, (NZ, NX, NY)).astype(np.float32)
    np.stack([vp, vs, rho], -1).astype("<f4").tofile(paths["media"])
    return paths
