"""The batched workload at full size: 48 Gaussian-blob volumes (seeds
1000 + v), axes drawn from 48..256 and 32..128 isovalues each, ds=2, 128
bins, marching cubes in compat mode, all pairs of all maps in one persistent
launch replayed from a captured HIP graph.

Every map has the shape every ISM has, and two graph replays give the same
maps.  The largest volume whose classify launch leaves a tail of cells out
(more than 8,388,480 voxels, MarchingCube.cpp:599-608) is checked against the
oracle: its map equals the single-volume pipeline's, vertex sets of two
isosurfaces and field values on sampled voxels equal the oracle's, and sampled
map entries equal the reference-order float NMI of the GPU's fields."""
import numpy as np
import pytest

import isoamd
import oracle

pytestmark = pytest.mark.gpu

NV, DS, H = 48, 2, 128


@pytest.fixture(scope="module")
def batch(ctx):
    rng = np.random.default_rng(7)
    dims = [tuple(int(v) for v in rng.integers(48, 257, 3)) for _ in range(NV)]
    ks = [int(v) for v in rng.integers(32, 129, NV)]
    imgs, isos, vols = [], [], []
    for v, (d, k) in enumerate(zip(dims, ks)):
        img = isoamd.gen_blobs(*d, nblobs=8, noise=0.01, seed=1000 + v)
        lo, hi = float(img.min()), float(img.max())
        isos.append(np.array([lo + (i + 1) * (hi - lo) / (k + 1) for i in range(k)], np.float32))
        imgs.append(img)
        vols.append(ctx.to_device(img))
    yield {"dims": dims, "ks": ks, "imgs": imgs, "isos": isos, "vols": vols}
    for d in vols:
        d.free()


def _run(ctx, b):
    return ctx.ism_batch(b["vols"], b["dims"], b["isos"], isoamd.EXTRACT_MC, DS, H,
                         isoamd.DF_LBVH, isoamd.COMPAT_ALL | isoamd.BATCH_GRAPH)


def test_c3_batch_maps_and_replays(ctx, batch):
    print("c3: batch capture + two replays", flush=True)
    first, st0 = _run(ctx, batch)
    again, st1 = _run(ctx, batch)
    third, st2 = _run(ctx, batch)
    assert st1.graph_replays >= 1 and st2.graph_replays == st1.graph_replays + 1
    assert st0.n_nonempty == sum(m.shape[0] for m, _ in first)
    for v, ((mp, ids), (mp2, ids2), (mp3, _)) in enumerate(zip(first, again, third)):
        m = mp.shape[0]
        assert 1 <= m <= batch["ks"][v]
        assert np.array_equal(mp, mp2) and np.array_equal(mp, mp3) and np.array_equal(ids, ids2)
        assert list(ids) == sorted(ids)
        assert np.all(np.diag(mp) == 1.0) and np.all((mp >= 0.0) & (mp <= 1.0))
    batch["maps"] = first


def test_c3_largest_truncated_volume_against_oracle(ctx, batch):
    if "maps" not in batch:
        batch["maps"], _ = _run(ctx, batch)
    big = [v for v in range(NV) if np.prod(batch["dims"][v]) > 128 * 65535]
    assert big, "the draw has volumes beyond the reference's classify launch"
    v = max(big, key=lambda q: int(np.prod(batch["dims"][q])))
    dims, isos, img = batch["dims"][v], batch["isos"][v], batch["imgs"][v]
    X, Y, Z = dims
    ncls = oracle.mc_ncls_compat(X, Y, Z)
    assert ncls < X * Y * Z
    print(f"c3: volume {v} {dims}, k={len(isos)}, {X * Y * Z - ncls} cells never classified",
          flush=True)
    smap, ids, st, fields = ctx.ism_volume(batch["vols"][v], dims, isos, isoamd.EXTRACT_MC, DS, H,
                                           isoamd.DF_LBVH, isoamd.COMPAT_ALL, want_fields=True)
    m = st.n_nonempty
    mp = smap.numpy()[: m * m].reshape(m, m)
    assert np.array_equal(mp, batch["maps"][v][0]) and np.array_equal(ids, batch["maps"][v][1])
    fv = fields.numpy()[:m]
    F = fv.shape[1]
    for r in (0, m // 2):
        iso = isos[ids[r]]
        out, n, stats = ctx.mc_extract(batch["vols"][v], dims, float(iso),
                                       isoamd.COMPAT_MC | isoamd.MC_SORTED)
        want, (active, total) = oracle.mc(img.ravel(), dims, iso, ncls=ncls)
        assert n == want.shape[0] and (stats.active_cells, stats.total_verts) == (active, total)
        assert np.array_equal(out.numpy()[:n], want)
        out.free()
        idx = np.sort(np.random.default_rng(r).choice(F, 2048, replace=False)).astype(np.int64)
        assert np.array_equal(fv[r][idx], oracle.df_tree(oracle.normalize(want, dims), dims, DS,
                                                          idx=idx))
    rng = np.random.default_rng(3)
    for _ in range(12):
        i, j = (int(q) for q in rng.integers(0, m, 2))
        h = oracle.hist_pair(fv[i], fv[j], H, *oracle.interval(fv[i]), *oracle.interval(fv[j]),
                             processed=oracle.hist_processed(F))
        assert np.float32(oracle.nmi_ref32(h, F)) == mp[i, j], (i, j)
    smap.free()
    fields.free()
