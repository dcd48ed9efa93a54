"""Restatements of the SgtHead tools that consume the raw SGT file (test
infrastructure: used only by tests/ as the checker of the device writers).

reformat_awp: SgtHead/src/reformat_awp.c:7-66 (and its MPI twin
reformat_awp_mpi.c:92-122, same per-point mapping).  The raw AWP file is read
as records of 6 * timesteps float32 (one point: XX YY ZZ XY XZ YZ series);
only complete records are processed (the fread loop, :38-62).  Per record:
for the Z source every value is first replaced by -2.0 * value (computed in
double, stored as float: :40-44); then the output record is
    XX <- YY, YY <- XX, ZZ <- ZZ, XY <- XY, XZ <- -1.0 * YZ, YZ <- -1.0 * XZ
(:46-61), i.e. the RWG component order and signs.
"""
from __future__ import annotations

import numpy as np


def reformat_awp(raw: bytes, timesteps: int, zsource: bool = False) -> bytes:
    a = np.frombuffer(raw, dtype="<f4")
    rec = 6 * timesteps
    nrec = a.size // rec
    r = a[: nrec * rec].reshape(nrec, 6, timesteps)
    if zsource:
        r = (-2.0 * r.astype(np.float64)).astype(np.float32)
    out = np.empty_like(r)
    out[:, 0] = r[:, 1]
    out[:, 1] = r[:, 0]
    out[:, 2] = r[:, 2]
    out[:, 3] = r[:, 3]
    out[:, 4] = (-1.0 * r[:, 5].astype(np.float64)).astype(np.float32)
    out[:, 5] = (-1.0 * r[:, 4].astype(np.float64)).astype(np.float32)
    return out.astype("<f4").tobytes()
