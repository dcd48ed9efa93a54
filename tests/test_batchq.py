"""Admission of runs to one GPU under a memory budget (host logic)."""
import pytest

from awpsgt.batchq import Need, admit, site_grids


def test_admit_packs_in_queue_order():
    gb = 10 ** 9
    needs = [Need(100 * gb, 110 * gb), Need(100 * gb, 105 * gb), Need(60 * gb, 70 * gb), Need(20 * gb, 21 * gb)]
    # 100+100+max(10,5) = 210 <= 288; +60 -> 270 + 10 = 280 <= 288; +20 -> 290 > 288
    assert admit(needs, 288 * gb) == [[0, 1, 2], [3]]
    assert admit(needs, 10 ** 15) == [[0, 1, 2, 3]]
    assert admit(needs, 115 * gb) == [[0], [1], [2, 3]]


def test_admit_counts_one_transient_per_wave():
    gb = 10 ** 9
    needs = [Need(50 * gb, 90 * gb), Need(50 * gb, 60 * gb)]
    # 100 steady + the larger transient (40) = 140
    assert admit(needs, 140 * gb) == [[0, 1]]
    assert admit(needs, 139 * gb) == [[0], [1]]


def test_admit_rejects_run_larger_than_budget():
    with pytest.raises(ValueError, match="budget"):
        admit([Need(10, 30)], 20)


def test_site_grids_interpolate_both_ends():
    g = site_grids(4, (400, 300, 150), (900, 700, 220))
    assert g[0] == (400, 300, 150) and g[-1] == (900, 700, 220)
    assert g[1] == (567, 433, 173)          # round(400 + 500/3), ...
    assert site_grids(1, (5, 6, 7), (9, 9, 9)) == [(5, 6, 7)]
