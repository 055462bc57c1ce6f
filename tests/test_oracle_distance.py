"""Pins oracle/distance.py (the `distance` module's loop,
runtime/src/DistanceEqualization.cpp) with answers fixed by geometry rather
than by the restatement itself: the line intersection of
binaryIsoValueSearch (:1421-1433) on hand-worked bounds, one refinement step
that must move every free isovalue to the midpoint of its neighbours when the
distances are the isovalue spacings, and the walk past empty isosurfaces of
approximation() (:84-186)."""
import numpy as np
import pytest

from oracle import distance

F32 = np.float32


def test_intersect_symmetric_bounds_give_the_midpoint():
    # first step of a pair: lower (iso_j, 0, S), upper (iso_k, S, 0)
    assert distance.intersect((F32(10), F32(0), F32(8)), (F32(18), F32(8), F32(0))) == F32(14)


def test_intersect_hand_worked():
    # left line through (8, 1) and (16, 5): 1 + (x - 8) / 2; right line through
    # (8, 7) and (16, 1): 7 - 3 (x - 8) / 4; they cross at x = 8 + 24/5 = 12.8
    got = distance.intersect((F32(8), F32(1), F32(7)), (F32(16), F32(5), F32(1)))
    assert got == F32(12.8)
    # left (10, 1) -> (18, 6), right (10, 5) -> (18, 2): cross at x = 14
    assert distance.intersect((F32(10), F32(1), F32(5)), (F32(18), F32(6), F32(2))) == F32(14)
    # a bound's left distance paired with the other bound's right one is a
    # different line: (8, 1) -> (16, 1) against (8, 7) -> (16, 5) cross at 32
    assert distance.intersect((F32(8), F32(1), F32(7)), (F32(16), F32(1), F32(5))) == F32(32)


def test_intersect_equal_bounds_is_nan():
    # 0/0: the reference leaves the isovalue unchanged (isnan check, :1435)
    assert np.isnan(distance.intersect((F32(5), F32(1), F32(2)), (F32(5), F32(1), F32(2))))


def _spacing(iso):
    iso = np.asarray(iso, np.float32)
    return np.array([F32(iso[j + 1] - iso[j]) for j in range(distance.NUM_DIST)], np.float32)


def test_one_step_moves_free_isovalues_to_neighbour_midpoints():
    # with d_j = iso[j+1] - iso[j], both distance lines of a pair have slope
    # +-1, so the first refinement step (odd isovalues free, even ones fixed)
    # puts every odd isovalue at the midpoint of its neighbours
    img = np.zeros((4, 4, 4), np.float32)
    img[0, 0, 0], img[-1, -1, -1] = 0.0, 40000.0
    out = distance.run(img, 1, max_itr=1, distances_fn=_spacing, cells_fn=lambda v: 1)
    est_tag, est_iso, est_dist = out[0]
    assert est_tag == "final" and np.array_equal(est_iso, distance.START)
    assert np.array_equal(est_dist, _spacing(distance.START))
    tag, iso, dist = out[1]
    assert tag == 0
    s = distance.START.astype(np.float64)
    for i in range(1, distance.NUM_DIST, 2):
        np.testing.assert_allclose(iso[i], 0.5 * (s[i - 1] + s[i + 1]), rtol=2e-7)
    for i in range(0, distance.NUM_DIST + 1, 2):
        assert iso[i] == distance.START[i]  # fixed set and end points untouched
    assert np.array_equal(dist, _spacing(iso))


def test_approximation_walks_past_empty_isosurfaces():
    # no cells below T: the first isovalue walks up in steps of
    # (iso[1] - iso[0]) / 16 and stops at the first value at or above T
    T = F32(137.65)
    img = np.zeros((4, 4, 4), np.float32)
    img[-1, -1, -1] = 40000.0
    out = distance.run(img, 1, max_itr=0, distances_fn=_spacing,
                       cells_fn=lambda v: 0 if v < T else 1)
    iso = out[0][1]
    incr = F32(F32(distance.START[1] - distance.START[0]) / F32(16))
    assert iso[0] >= T and F32(iso[0] - incr) < T
    steps = round(float((iso[0] - distance.START[0]) / incr))
    assert steps == int(np.ceil((137.65 - float(distance.START[0])) / float(incr)))
    assert np.array_equal(iso[1:], distance.START[1:])
