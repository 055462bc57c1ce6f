"""Volume ingest (x-fastest file -> z-fastest image, restricted extremes) and
the 65536-bin isovalue equalisation: oracle pinned on the CPU, HIP bit-exact."""
import numpy as np
import pytest

import isoamd
import oracle


def test_oracle_ingest_layout_and_extremes():
    X, Y, Z = 7, 5, 3
    file = np.arange(X * Y * Z, dtype=np.float32)  # value = x + y*X + z*X*Y
    img, lo, hi = oracle.ingest(file, (X, Y, Z), 4.0, 50.0)
    x, y, z = np.meshgrid(np.arange(X), np.arange(Y), np.arange(Z), indexing="ij")
    assert np.array_equal(img, (x + y * X + z * X * Y).astype(np.float32))
    assert lo == 4.0 and hi == 50.0
    _, lo2, hi2 = oracle.ingest(file, (X, Y, Z), 1e6, -1e6)  # empty window keeps the seeds
    assert lo2 == 100000.0 and hi2 == -100000.0


def test_oracle_equalisation_closed_forms():
    # uniform values: equalised isovalues split the range evenly
    v = (np.arange(65536 * 4, dtype=np.float64) / (65536 * 4)).astype(np.float32)
    iso = oracle.equalized_isovalues(v, float(v.min()), float(v.max()), 8)
    assert np.all(np.diff(iso) > 0)
    np.testing.assert_allclose(iso, (np.arange(8) + 0.5) / 8, atol=2e-4)
    # a skewed distribution (t^3): the bins split into equal-count groups at the
    # quantiles q(i/4) = (i/4)^3, and each isovalue is the middle bin of its group
    w = (np.linspace(0, 1, 200000) ** 3).astype(np.float32)
    iw = oracle.equalized_isovalues(w, float(w.min()), float(w.max()), 4)
    edges = (np.arange(5) / 4.0) ** 3
    np.testing.assert_allclose(iw, (edges[:-1] + edges[1:]) / 2, atol=5e-4)


def test_oracle_equalisation_float_loop_bound():
    # the reference loops `i < dataSize` with a float dataSize: for n = 2^24 + 3
    # float(n) == 2^24 + 4, so every voxel counts; for n = 36902400 the last two
    # voxels (whose float index rounds up to n) are skipped
    n = 36902400
    a = np.zeros(n, np.float32)
    a[-2:] = 1.0  # only the skipped voxels hold the top value
    a[0] = 0.5
    iso = oracle.equalized_isovalues(a, 0.0, 1.0, 2)
    b = a.copy()
    b[-2:] = 0.0
    assert np.array_equal(iso, oracle.equalized_isovalues(b, 0.0, 1.0, 2))


@pytest.mark.gpu
@pytest.mark.parametrize("dims", [(64, 48, 40), (600, 248, 248), (33, 1, 17)])
def test_ingest_matches_oracle(ctx, dims):
    X, Y, Z = dims
    rng = np.random.default_rng(X)
    file = (rng.random(X * Y * Z) * 1000).astype(np.float32)
    cmin, cmax = 100.0, 900.0
    img, lo, hi = ctx.volume_ingest(ctx.to_device(file), dims, cmin, cmax)
    want, wlo, whi = oracle.ingest(file, dims, cmin, cmax)
    assert np.array_equal(img.numpy(), want) and (lo, hi) == (wlo, whi)


@pytest.mark.gpu
@pytest.mark.parametrize("k", [16, 256])
def test_equalized_isovalues_match_oracle(ctx, k):
    dims = (600, 248, 248)
    img = isoamd.gen_front(*dims, seed=2)
    d = ctx.to_device(img)
    lo, hi = float(img.min()), float(img.max())
    got = ctx.equalized_isovalues(d, lo, hi, k)
    want = oracle.equalized_isovalues(img, lo, hi, k)
    assert np.array_equal(got, want)
