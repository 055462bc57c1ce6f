"""Representative selection (host, no GPU): hand-derived small cases, the
oracle restatement of SimilarityMap.cpp on random maps, and the ascending-id
form used for distancefield-<n>.df naming."""
import numpy as np
import pytest

import isoamd
import oracle


def test_two_surfaces_closed_form():
    s = np.float32(0.5)
    smap = np.array([[1, s], [s, 1]], np.float32)
    sel, rows, pri = isoamd.select_representatives_ordered(smap, 2)
    # both means 0.75, first maximum wins: priority 2*0.75; the other set {1}
    # has priority 1, divided by 1+0.5 after the first pick
    assert list(sel) == [0, 1] and list(rows) == [0, 1]
    assert pri[0] == np.float32(1.5) and pri[1] == np.float32(1.0) / np.float32(1.5)


def test_three_surfaces_hand_derived():
    smap = np.array([[1, .8, .1], [.8, 1, .2], [.1, .2, 1]], np.float32)
    sel, rows, pri = isoamd.select_representatives_ordered(smap, 3, ids=[10, 20, 30])
    # means .6333/.6667/.4333 -> row 1 first (priority 3*mean = 2); {0} and {2}
    # get priority 1; after the pick: 1/1.8 and 1/1.2 -> row 2; then row 0 with
    # (1/1.8)/1.1
    assert list(rows) == [1, 2, 0] and list(sel) == [20, 30, 10]
    np.testing.assert_allclose(pri, [2.0, 1 / 1.2, 1 / 1.8 / 1.1], rtol=1e-6)


@pytest.mark.parametrize("m,cnt,seed", [(5, 3, 0), (17, 16, 1), (40, 16, 2), (64, 16, 3), (9, 20, 4)])
def test_ordered_matches_oracle(m, cnt, seed):
    rng = np.random.default_rng(seed)
    a = rng.random((m, m)).astype(np.float32)
    smap = ((a + a.T) * np.float32(0.5)).astype(np.float32)
    np.fill_diagonal(smap, 1.0)
    ids = np.sort(rng.choice(4 * m, m, replace=False)).astype(np.int32)
    sel, rows, pri = isoamd.select_representatives_ordered(smap, cnt, ids=ids)
    want = oracle.representatives(smap, cnt, ids)
    assert [w[0] for w in want] == list(sel)
    assert [w[1] for w in want] == list(rows)
    assert np.array_equal(np.array([w[2] for w in want], np.float32), pri)
    asc, apri = isoamd.select_representatives(smap, cnt, ids=ids)
    order = np.argsort(sel)
    assert np.array_equal(asc, sel[order]) and np.array_equal(apri, pri[order])
