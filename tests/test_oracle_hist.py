"""Pins the oracle's histogram / NMI restatement (JointHistogram.cu:24-162) with
closed forms and invariants, independent of the oracle's own formulas."""
import math

import numpy as np
import pytest

import oracle


def test_bins_on_known_edges():
    # range [0,1], H=4: step 0.25; 0.25 exactly is the start of bin 1;
    # the maximum is clamped into bin H-1 (findBucket_d upper clamp).
    f = np.array([0.0, 0.1, 0.25, 0.5, 0.74, 0.75, 1.0], np.float32)
    h = oracle.hist_pair(f, f, 4, 0.0, 1.0, 0.0, 1.0)
    assert np.array_equal(np.diag(h), [2, 1, 2, 2])
    assert h.sum() == f.size and np.count_nonzero(h - np.diag(np.diag(h))) == 0


def test_pair_range_is_union_of_intervals():
    # field a in [0,1], field b in [2,4]: common range [0,4], step 0.5 (H=8)
    a = np.array([0.0, 1.0], np.float32)
    b = np.array([2.0, 4.0], np.float32)
    h = oracle.hist_pair(a, b, 8, 0.0, 1.0, 2.0, 4.0)
    assert h[0, 4] == 1 and h[2, 7] == 1 and h.sum() == 2


def test_zero_step_puts_everything_in_bin0():
    f = np.full(100, 3.5, np.float32)
    h = oracle.hist_pair(f, f, 16, 3.5, 3.5, 3.5, 3.5)
    assert h[0, 0] == 100 and h.sum() == 100


def test_single_bin():
    rng = np.random.default_rng(0)
    f = rng.random(1000).astype(np.float32)
    h = oracle.hist_pair(f, f[::-1].copy(), 1, float(f.min()), float(f.max()),
                         float(f.min()), float(f.max()))
    assert h.shape == (1, 1) and h[0, 0] == 1000


def test_marginals_match_one_dimensional_histograms():
    rng = np.random.default_rng(1)
    a = rng.random(5000).astype(np.float32)
    b = (rng.random(5000) * 2).astype(np.float32)
    H = 32
    mn, mx = 0.0, float(max(a.max(), b.max()))
    h = oracle.hist_pair(a, b, H, float(a.min()), float(a.max()), float(b.min()), float(b.max()))
    lo = float(min(a.min(), b.min()))
    step = np.float32((np.float32(mx) - np.float32(lo)) / np.float32(H))
    ra = np.minimum(((a - np.float32(lo)) / step).astype(np.int32), H - 1)
    rb = np.minimum(((b - np.float32(lo)) / step).astype(np.int32), H - 1)
    assert np.array_equal(h.sum(1), np.bincount(ra, minlength=H))
    assert np.array_equal(h.sum(0), np.bincount(rb, minlength=H))


def test_nmi_identical_fields_is_one_and_constant_is_zero():
    rng = np.random.default_rng(2)
    f = rng.random(10000).astype(np.float32)
    lo, hi = oracle.interval(f)
    h = oracle.hist_pair(f, f, 64, lo, hi, lo, hi)
    assert oracle.nmi_ref32(h, f.size) == 1.0
    assert oracle.nmi_f64(h, f.size) == 1.0
    c = np.full(100, 2.0, np.float32)
    hc = oracle.hist_pair(c, c, 8, 2.0, 2.0, 2.0, 2.0)
    assert oracle.nmi_ref32(hc, 100) == 0.0 and oracle.nmi_f64(hc, 100) == 0.0


def test_nmi_closed_form_2x2():
    # joint counts [[a, b], [c, d]]: entropies from the textbook definition
    for counts in ([[30, 10], [5, 55]], [[25, 25], [25, 25]], [[50, 0], [0, 50]], [[1, 2], [3, 4]]):
        h = np.array(counts, np.uint32)
        N = int(h.sum())
        p = h / N
        px, py = p.sum(1), p.sum(0)
        ent = lambda q: -sum(x * math.log(x) for x in q.ravel() if x > 0)
        hx, hy, hxy = ent(px), ent(py), ent(p)
        want = 2 * (hx + hy - hxy) / (hx + hy)
        assert oracle.nmi_f64(h, N) == pytest.approx(want, rel=1e-12, abs=1e-15)
        assert oracle.nmi_ref32(h, N) == pytest.approx(want, rel=1e-5, abs=1e-6)


def test_nmi_independent_fields_near_zero():
    rng = np.random.default_rng(3)
    a = rng.random(200000).astype(np.float32)
    b = rng.random(200000).astype(np.float32)
    h = oracle.hist_pair(a, b, 8, float(a.min()), float(a.max()), float(b.min()), float(b.max()))
    assert 0 <= oracle.nmi_f64(h, a.size) < 1e-3


def test_chunk_quirk_counts():
    # JointHistogram.cu:377-393: the last 1M chunk is skipped iff F % 1M == 0
    assert oracle.hist_processed(576600) == 576600
    assert oracle.hist_processed(2097152) == 2097152
    assert oracle.hist_processed(1000000) == 0
    assert oracle.hist_processed(2000000) == 1000000


def test_chunk_quirk_gives_all_ones_map():
    # F = 1,000,000: no voxel is binned, every entry is 2 ln N / 2 ln N = 1
    rng = np.random.default_rng(4)
    fields = rng.random((3, 1000000)).astype(np.float32)
    m32, m64 = oracle.ism(fields, 16, compat_chunks=True)
    assert np.all(m32 == 1.0) and np.all(m64 == 1.0)


def test_ism_diagonal_and_symmetry():
    rng = np.random.default_rng(5)
    base = rng.random(20000).astype(np.float32)
    fields = np.stack([base + 0.1 * k * rng.random(20000).astype(np.float32) for k in range(5)])
    m32, m64 = oracle.ism(fields, 32)
    assert np.all(np.diag(m32) == 1.0) and np.all(np.diag(m64) == 1.0)
    assert np.allclose(m64, m64.T, rtol=1e-12)
    # reference float order stays within the ISM tolerance of the exact value here
    assert np.allclose(m32, m64, rtol=1e-5, atol=1e-6)


def test_interval_uses_flt_min_for_max():
    lo, hi = oracle.interval(np.zeros(10, np.float32))
    assert lo == 0.0 and hi == np.finfo(np.float32).tiny


# ---- reference-order float NMI: pinned by exact rational arithmetic --------
# calculateMutualInformation (JointHistogram.cu:24-60) is a fixed sequence of
# float roundings: fmaf chains hXY (bins row by row), hX (column sums), hY
# (row sums), then h/N + logf(N) and 2(hX+hY-hXY)/(hX+hY).  Below, every
# rounding is done from exact rationals with round-half-even to float32, so a
# swapped order, an unfused multiply-add or a wrong final formula in the C
# oracle shows up as a different bit pattern.

from fractions import Fraction  # noqa: E402


def _f32(q):
    """float32 nearest the rational q, ties to even."""
    q = Fraction(q)
    if q == 0:
        return np.float32(0.0)
    x = np.float32(float(q))
    cands = [np.nextafter(x, np.float32(-np.inf)), x, np.nextafter(x, np.float32(np.inf))]
    return min(cands, key=lambda c: (abs(Fraction(float(c)) - q), int(c.view(np.uint32)) & 1))


def _logf(c):
    return np.float32(math.log(float(np.float32(c))))


def _chain(counts):
    acc = np.float32(0.0)
    for c in counts:
        if c:
            p = np.float32(c)
            acc = _f32(-Fraction(float(p)) * Fraction(float(_logf(p))) + Fraction(float(acc)))
    return acc


def _nmi_exact_rounding(hist, N, order="row"):
    H = hist.shape[0]
    cells = hist.ravel() if order == "row" else hist.T.ravel()
    hxy = _chain(cells)
    hx = _chain(hist.sum(0))  # colSums
    hy = _chain(hist.sum(1))  # rowSums
    fN = np.float32(N)
    lN = _logf(fN)
    fin = lambda h: _f32(Fraction(float(_f32(Fraction(float(h)) / Fraction(float(fN))))) +
                         Fraction(float(lN)))
    hXY, hX, hY = fin(hxy), fin(hx), fin(hy)
    s = _f32(Fraction(float(hX)) + Fraction(float(hY)))
    i_xy = _f32(Fraction(float(s)) - Fraction(float(hXY)))
    num = _f32(2 * Fraction(float(i_xy)))
    if s == 0:  # 0/0 is NaN -> 0; x/0 is an infinity
        return np.float32(0.0) if num == 0 else np.float32(np.inf) * np.sign(num)
    v = _f32(Fraction(float(num)) / Fraction(float(s)))
    return np.float32(0.0) if v != v else v


def test_nmi_ref32_equals_exact_rounding_sequence():
    rng = np.random.default_rng(8)
    order_mattered = 0
    for t in range(40):
        H = int(rng.integers(2, 6))
        hist = rng.integers(0, 4000, (H, H)).astype(np.uint32)
        hist[rng.random((H, H)) < 0.3] = 0
        N = int(hist.sum())
        if N == 0:
            continue
        want = _nmi_exact_rounding(hist, N, "row")
        got = np.float32(oracle.nmi_ref32(hist, N))
        assert got.view(np.uint32) == want.view(np.uint32), (t, hist, got, want)
        order_mattered += _nmi_exact_rounding(hist, N, "col") != want
    assert order_mattered > 0  # the draw includes cases where the order shows


def test_nmi_ref32_within_float_error_of_exact_value():
    # the float chains stay within their rounding-error bound of the exact NMI
    # (from the double evaluation): no term dropped or counted twice
    rng = np.random.default_rng(3)
    for _ in range(20):
        H = int(rng.integers(4, 64))
        hist = rng.integers(0, 50, (H, H)).astype(np.uint32) * (rng.random((H, H)) < 0.5)
        hist = hist.astype(np.uint32)
        N = int(hist.sum())
        e32 = oracle.nmi_ref32(hist, N)
        e64 = oracle.nmi_f64(hist, N)
        hx = math.log(N) - sum(c * math.log(c) for c in hist.sum(0) if c) / N
        # each accumulator carries at most H*H roundings of size 2^-24 |sum|
        bound = 4 * (H * H + 4) * 2.0 ** -24 * (math.log(N) + 1) / max(hx, 1e-3)
        assert abs(e32 - e64) <= bound * max(abs(e64), 1e-3) + 1e-6


def test_logf_cr_has_one_answer_for_every_count_below_16200282():
    # (float)log((double)c) is the correctly rounded logf for every integer
    # count c < 16,200,282: no exact logarithm sits within 2^-48 (relative) of
    # a float rounding boundary, so any double log rounds the same way
    n, first = oracle.logf_cr_margin(1.0, 16200281.0)
    assert n == 0, first
    n, first = oracle.logf_cr_margin(1.0, 2.0 ** 32)
    assert first == 16200282.0 and n == 9


def test_nmi_ref32_sensitivity_to_a_one_ulp_logf_is_small_on_field_pairs():
    # another logf within 1 ulp (as CUDA documents its own) moves reference-order
    # NMI entries of correlated distance-field-like pairs by far less than 1e-5
    rng = np.random.default_rng(9)
    F = 40000
    base = rng.random(F).astype(np.float32)
    a = np.abs(base - 0.3).astype(np.float32)
    b = np.abs(base - 0.35 + 0.01 * rng.standard_normal(F).astype(np.float32)).astype(np.float32)
    h = oracle.hist_pair(a, b, 128, *oracle.interval(a), *oracle.interval(b))
    v0 = oracle.nmi_ref32(h, F)
    worst = max(abs(oracle.nmi_ref32_ulp(h, F, s) - v0) / v0 for s in range(1, 9))
    assert worst < 1e-5
