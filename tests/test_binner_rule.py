"""The histogram kernels' fast binning rule, replayed in IEEE float32 on the
CPU: q = (v - mn) * RN(1 / step) with step = RN((mx - mn) / H).  Whenever
|frac(q) - 1/2| <= 1/2 - H 2^-21 (the rule's "not near an integer"), trunc(q)
must be the reference's bin min(trunc(RN((v - mn) / step)), H - 1)
(JointHistogram.cu:63-75, :127-129) and below H without a clamp.  Values at
the bin edges, at mx, and random values over many ranges and bin counts;
float32 numpy arithmetic rounds like the GPU's (correctly rounded division).
"""
import numpy as np

F = np.float32


def reference_bin(v, mn, step, H):
    a = (v - mn).astype(F)
    return np.minimum((a / step).astype(F).astype(np.int64), H - 1)


def fast_rule(v, mn, rstep, H):
    a = (v - mn).astype(F)
    q = (a * rstep).astype(F)
    fr = (q - np.floor(q)).astype(F)
    margin = F(H * 2.0 ** -21)
    ok = np.abs((fr - F(0.5)).astype(F)) <= F(F(0.5) - margin)
    return q.astype(np.int64), ok


def _check(v, mn, mx, H):
    v = v.astype(F)
    step = F((F(mx - mn)) / F(H))
    if step == 0:
        return 0
    rstep = F(F(1.0) / step)
    ref = reference_bin(v, F(mn), step, H)
    k, ok = fast_rule(v, F(mn), rstep, H)
    assert np.all(k[ok] == ref[ok]), (mn, mx, H)
    assert np.all(k[ok] <= H - 1)
    return int((~ok).sum())


def test_fast_rule_matches_reference_bins():
    rng = np.random.default_rng(7)
    slow = total = 0
    for H in (1, 2, 7, 64, 100, 128, 255, 256):
        for _ in range(40):
            mn = F(rng.uniform(-50, 50) * 10.0 ** rng.integers(-3, 4))
            mx = F(mn + abs(rng.normal()) * 10.0 ** rng.integers(-4, 4))
            if not mx > mn:
                continue
            step = F((F(mx - mn)) / F(H))
            ks = np.arange(H + 1, dtype=np.float64)
            edges = (mn + ks * np.float64(step)).astype(F)
            near = np.concatenate([edges, np.nextafter(edges, F(np.inf)),
                                   np.nextafter(edges, F(-np.inf))])
            near = near[(near >= mn) & (near <= mx)]
            v = np.concatenate([near, rng.uniform(mn, mx, 20000).astype(F), [F(mx), F(mn)]])
            v = np.clip(v, mn, mx).astype(F)
            slow += _check(v, mn, mx, H)
            total += v.size
    # the exact path stays rare for random values
    assert slow < total


def test_top_value_takes_exact_path():
    # v = mx gives q at or just past H: never on the fast path
    for H in (1, 3, 128, 256):
        for mn, mx in ((0.0, 1.0), (0.1, 7.3), (-3.0, 1e-3), (1e-7, 2e-7)):
            v = np.array([mx], F)
            step = F((F(mx) - F(mn)) / F(H))
            rstep = F(F(1.0) / step)
            k, ok = fast_rule(v, F(mn), rstep, H)
            assert (not ok[0]) or k[0] <= H - 1
