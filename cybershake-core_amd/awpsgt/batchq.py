"""Admission of SGT runs to one GPU under a device-memory budget.

A queue of runs (for example the X and Y components of several sites) is
split into waves: each wave is a set of runs that are resident together and
advanced as one batch (awpsgt.Batch).  Runs are taken in queue order; a run
joins the current wave while the wave's steady bytes plus the largest
set-up transient of its runs (the media file and temporaries, present for one
run at a time while its media are set) stay within the budget; otherwise the
wave is closed and the run starts the next one.  Byte counts come from
awp_estimate_bytes (the same arithmetic as awp_create's allocation).

The reference has no counterpart: it runs one component per MPI job
(AWP-ODC-SGT/utils/awp_odc_wrapper.sh:14).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Sequence, Tuple


@dataclass
class Need:
    """Device bytes of one queued run."""
    steady: int
    peak: int

    @property
    def transient(self) -> int:
        return self.peak - self.steady


def admit(needs: Sequence[Need], budget: int) -> List[List[int]]:
    """Waves of run indices (queue order kept) under ``budget`` bytes.

    Raises ValueError if a single run does not fit on its own."""
    waves: List[List[int]] = []
    cur: List[int] = []
    steady = 0
    trans = 0
    for k, n in enumerate(needs):
        if n.peak > budget:
            raise ValueError("run %d needs %.1f GB, more than the %.1f GB budget"
                             % (k, n.peak / 1e9, budget / 1e9))
        if cur and steady + n.steady + max(trans, n.transient) > budget:
            waves.append(cur)
            cur, steady, trans = [], 0, 0
        cur.append(k)
        steady += n.steady
        trans = max(trans, n.transient)
    if cur:
        waves.append(cur)
    return waves


def site_grids(nsites: int, lo: Tuple[int, int, int], hi: Tuple[int, int, int]) -> List[Tuple[int, int, int]]:
    """Grids linearly interpolated from ``lo`` to ``hi`` over ``nsites``
    sites: site s gets round(lo + (hi - lo) * s / (nsites - 1)) per axis."""
    out = []
    for s in range(nsites):
        f = s / max(nsites - 1, 1)
        out.append(tuple(int(round(a + (b - a) * f)) for a, b in zip(lo, hi)))
    return out
