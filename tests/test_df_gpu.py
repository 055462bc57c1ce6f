"""HIP distance fields: brute force against the oracle bit for bit, and the LBVH
against the brute force bit for bit (the exact-minimum property)."""
import numpy as np
import pytest

import isoamd
import oracle

pytestmark = pytest.mark.gpu


def _surface(img, iso, compat=True):
    X, Y, Z = img.shape
    want, _ = oracle.mc(img.ravel(), (X, Y, Z), np.float32(iso), ncls=oracle.mc_ncls_compat(X, Y, Z))
    return want


@pytest.mark.parametrize("dims,ds,seed", [((32, 32, 32), 1, 1), ((40, 28, 20), 2, 3), ((24, 17, 33), 3, 5)])
def test_brute_and_lbvh_match_oracle(ctx, dims, ds, seed):
    print(f"df parity dims={dims} ds={ds}", flush=True)
    img = isoamd.gen_blobs(*dims, seed=seed)
    iso = float(img.min() + 0.4 * (img.max() - img.min()))
    pts = _surface(img, iso)
    assert pts.shape[0] > 10
    d = ctx.to_device(pts)
    want = oracle.df_brute(oracle.normalize(pts, dims), dims, ds)
    brute = ctx.df_points(d, pts.shape[0], dims, ds, isoamd.DF_BRUTE).numpy()
    lbvh = ctx.df_points(d, pts.shape[0], dims, ds, isoamd.DF_LBVH).numpy()
    assert np.array_equal(brute, want)
    assert np.array_equal(lbvh, want)


def test_int_cells_path(ctx):
    dims = (36, 30, 22)
    img = isoamd.gen_blobs(*dims, seed=8)
    iso = float(img.mean())
    cells = oracle.approx(img, np.float32(iso))
    d = ctx.to_device(cells)
    want = oracle.df_brute(oracle.normalize(cells, dims), dims, 2)
    lbvh = ctx.df_points(d, cells.shape[0], dims, 2, isoamd.DF_LBVH).numpy()
    assert np.array_equal(lbvh, want)


@pytest.mark.parametrize("n", [1, 2, 3, 9, 17, 32, 33, 64, 65, 100])
def test_tiny_point_sets(ctx, n):
    dims = (20, 16, 12)
    rng = np.random.default_rng(n)
    pts = (rng.random((n, 3)) * dims).astype(np.float32)
    d = ctx.to_device(pts)
    want = oracle.df_brute(oracle.normalize(pts, dims), dims, 1)
    for m in (isoamd.DF_BRUTE, isoamd.DF_LBVH):
        got = ctx.df_points(d, n, dims, 1, m).numpy()
        assert np.array_equal(got, want), (n, m)


def test_duplicate_points_and_equal_codes(ctx):
    # many identical points and points sharing Morton codes
    dims = (64, 64, 64)
    rng = np.random.default_rng(7)
    base = (rng.random((50, 3)) * 64).astype(np.float32)
    pts = np.concatenate([np.repeat(base, 40, axis=0), base + np.float32(1e-3)]).astype(np.float32)
    d = ctx.to_device(pts)
    want = oracle.df_brute(oracle.normalize(pts, dims), dims, 4)
    for m in (isoamd.DF_BRUTE, isoamd.DF_LBVH):
        assert np.array_equal(ctx.df_points(d, pts.shape[0], dims, 4, m).numpy(), want)


def test_lbvh_equals_brute_at_scale(ctx):
    dims = (128, 96, 80)
    img = isoamd.gen_sines(*dims, seed=4)
    out, n, _ = ctx.mc_extract(ctx.to_device(img), dims, 0.2, isoamd.COMPAT_MC)
    assert n > 10000
    brute = ctx.df_points(out, n, dims, 2, isoamd.DF_BRUTE).numpy()
    lbvh = ctx.df_points(out, n, dims, 2, isoamd.DF_LBVH).numpy()
    assert np.array_equal(brute, lbvh)
    # spot-check against the oracle on a sample subset
    pts = out.numpy()[:n]
    idx = np.random.default_rng(0).choice(brute.size, 2000, replace=False)
    ref = oracle.df_brute_subset(oracle.normalize(pts, dims), dims, 2, idx)
    assert np.array_equal(brute[idx], ref)


@pytest.mark.parametrize("pool", ["1", "0"])
@pytest.mark.parametrize("leaf,spl", [("64", "4"), ("48", "2"), ("64", "2"), ("64", "1"),
                                      ("48", "1"), ("32", "2"), ("24", "4"), ("16", "2"),
                                      ("8", "1")])
def test_every_lbvh_variant_equals_brute(ctx, monkeypatch, leaf, spl, pool):
    # B, the most points per LBVH leaf, and samples per lane: each variant the
    # autotuner may keep (and the extremes), on a pool of isosurfaces of very
    # different sizes (one ISM); leaves cut from the extraction's runs
    # (pool 1, the pipeline's default) and from the sorted points (pool 0)
    print(f"lbvh leaf size {leaf} samples per lane {spl} pool leaves {pool}", flush=True)
    monkeypatch.setenv("ISO_DF_LEAF", leaf)
    monkeypatch.setenv("ISO_DF_SPL", spl)
    monkeypatch.setenv("ISO_DF_POOL_LEAVES", pool)
    dims, ds = (56, 44, 40), 2
    img = isoamd.gen_blobs(*dims, seed=12)
    lo, hi = float(img.min()), float(img.max())
    isos = np.array([lo + f * (hi - lo) for f in (0.02, 0.1, 0.3, 0.5, 0.7, 0.9, 0.98)], np.float32)
    d = ctx.to_device(img)
    _, _, s1, f1 = ctx.ism_volume(d, dims, isos, isoamd.EXTRACT_MC, ds, 64, isoamd.DF_LBVH,
                                  want_fields=True)
    _, _, s2, f2 = ctx.ism_volume(d, dims, isos, isoamd.EXTRACT_MC, ds, 64, isoamd.DF_BRUTE,
                                  want_fields=True)
    m = s1.n_nonempty
    assert m == s2.n_nonempty and m >= 5
    assert np.array_equal(f1.numpy()[:m], f2.numpy()[:m])


@pytest.mark.parametrize("spl", ["1", "2", "4"])
def test_lbvh_work_queue_equals_brute(ctx, monkeypatch, spl):
    # enough bricks for the walk's ticket queue (about 65k bricks and more on
    # a 256-CU part: whole columns dealt to the XCDs in runs of an eighth of a
    # set, a ragged last round, then half and quarter columns from the shared
    # counter); every field of every set against brute force
    print(f"lbvh work queue, samples per lane {spl}", flush=True)
    monkeypatch.setenv("ISO_DF_LEAF", "64")
    monkeypatch.setenv("ISO_DF_SPL", spl)
    dims, ds = (200, 120, 120), 2
    img = isoamd.gen_blobs(*dims, nblobs=24, seed=21)
    lo, hi = float(img.min()), float(img.max())
    isos = np.linspace(lo + 0.02 * (hi - lo), hi - 0.02 * (hi - lo), 48).astype(np.float32)
    d = ctx.to_device(img)
    _, _, s1, f1 = ctx.ism_volume(d, dims, isos, isoamd.EXTRACT_MC, ds, 64, isoamd.DF_LBVH,
                                  want_fields=True)
    _, _, s2, f2 = ctx.ism_volume(d, dims, isos, isoamd.EXTRACT_MC, ds, 64, isoamd.DF_BRUTE,
                                  want_fields=True)
    m = s1.n_nonempty
    nx, ny, nz = (n // ds for n in dims)
    s = int(spl)
    bz = (nz + 4 * s - 1) // (4 * s)
    bricks = m * ((nx + 3) // 4) * ((ny + 3) // 4) * bz
    print(f"sets {m}, bricks {bricks}", flush=True)
    assert m == s2.n_nonempty and m >= 40
    # 8192 waves: the last 8192 columns' worth go out in half and quarter
    # columns, the rest (here over a third) as whole columns
    assert bricks > 12288 * bz
    assert np.array_equal(f1.numpy()[:m], f2.numpy()[:m])


@pytest.mark.parametrize("items", ["4", "8"])
def test_sort_tile_sizes_give_identical_fields(ctx, monkeypatch, items):
    # the LBVH's keyed radix sort at 1024 x items keys per tile against the
    # default 16: a pool of several tiles with a ragged last one
    print(f"radix sort items per lane {items}", flush=True)
    monkeypatch.setenv("ISO_DF_POOL_LEAVES", "0")  # the point sort
    dims, ds = (56, 44, 40), 2
    img = isoamd.gen_blobs(*dims, seed=14)
    lo, hi = float(img.min()), float(img.max())
    isos = np.array([lo + f * (hi - lo) for f in (0.05, 0.3, 0.5, 0.7, 0.95)], np.float32)
    d = ctx.to_device(img)
    _, _, s1, f1 = ctx.ism_volume(d, dims, isos, isoamd.EXTRACT_MC, ds, 64, isoamd.DF_LBVH,
                                  want_fields=True)
    want = f1.numpy()[: s1.n_nonempty].copy()
    assert s1.total_points > 3 * 16384 and s1.total_points % 1024
    monkeypatch.setenv("ISO_SORT_ITEMS", items)
    _, _, s2, f2 = ctx.ism_volume(d, dims, isos, isoamd.EXTRACT_MC, ds, 64, isoamd.DF_LBVH,
                                  want_fields=True)
    assert s2.n_nonempty == s1.n_nonempty
    assert np.array_equal(f2.numpy()[: s2.n_nonempty], want)


def test_empty_point_set_is_an_error(ctx):
    d = ctx.empty((1, 3), np.float32)
    with pytest.raises(isoamd.IsoError) as e:
        ctx.df_points(d, 0, (8, 8, 8), 1)
    assert e.value.status == isoamd.ERR_EMPTY
