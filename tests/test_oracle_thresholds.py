"""The oracle's solid voxelizer at its thresholds: hand-derived known answers
(tests/kats.py) for the |n.x| skip, the |t| on-edge test and the top-left rule
of cpu_voxelizer.cpp:196-238,265, in both oracle modes and both layouts."""
import numpy as np
import pytest

from tests import kats

CASES = kats.all_cases()


@pytest.fixture(scope="module")
def info(oracle):
    return oracle.make_info(kats.BOX_MIN, kats.BOX_MAX, kats.GRID, 1)


@pytest.mark.parametrize("case", CASES, ids=[c.name for c in CASES])
def test_oracle_threshold_known_answers(oracle, info, case):
    for mode in (oracle.MODE_REF, oracle.MODE_FAST):
        for morton in (False, True):
            table, offrow, _ = oracle.solid(info, case.tri, morton=morton, mode=mode)
            assert offrow == 0
            occ = oracle.occupancy(table, kats.GRID, morton=morton)
            exp = case.occupancy()
            bad = np.argwhere(occ != exp)
            assert bad.size == 0, (mode, morton, bad[:8].tolist())


def test_cases_cover_both_sides_of_each_threshold():
    names = [c.name for c in CASES]
    nx = {c.name.split("-")[0]: bool(c.cols) for c in kats.normal_threshold_cases()}
    assert list(nx.values()) == [False, True]  # skipped below 1e-6, voxelized above
    t = {(c.name.split("-")[1], c.name.split("-")[2]): sorted(c.cols) for c in kats.edge_threshold_cases()}
    assert t == {("m8796093", "a"): [], ("m8796094", "a"): [(1, 0)],
                 ("m8796093", "b"): [(1, 0), (1, 1)], ("m8796094", "b"): [(1, 1)]}
    # the edge triangles keep columns on included edges only (T1: C->A; T2: both slanted
    # edge A'->C' and the horizontal top edge B'->A'), and every vertex order agrees
    ties = {c.name: c.cols for c in kats.edge_tie_cases()}
    t1 = ties["T1-abc"]
    assert (0, 1) in t1 and (1, 3) in t1 and (2, 5) in t1            # on C -> A: kept
    assert not any((y, 0) in t1 for y in range(6))                   # on A -> B: dropped
    assert (5, 1) not in t1 and (4, 3) not in t1 and (3, 5) not in t1  # on B -> C: dropped
    t2 = ties["T2-abc"]
    assert all((y, 7) in t2 for y in range(6))                       # on B' -> A': kept
    assert (3, 2) not in t2 and (4, 4) not in t2 and (5, 6) not in t2  # on C' -> B': dropped
    assert len({tuple(sorted(v.items())) for k, v in ties.items() if k.startswith("T1")}) == 1
    assert len({tuple(sorted(v.items())) for k, v in ties.items() if k.startswith("T2")}) == 1
    assert len(names) == len(set(names))
