"""iso_ism_batch: the ISMs of several volumes of uneven size in one call must
equal, map by map and bit for bit, what iso_ism_volume returns per volume
(which the ISM tests pin against the oracle)."""
import numpy as np
import pytest

import isoamd

pytestmark = pytest.mark.gpu


def _isos(img, k):
    lo, hi = float(img.min()), float(img.max())
    return np.array([lo + (i + 1) * (hi - lo) / (k + 1) for i in range(k)], np.float32)


def _per_volume(ctx, vols, dims_list, isos_list, extractor, ds, H, flags):
    out = []
    for v, dims, isos in zip(vols, dims_list, isos_list):
        smap, ids, st, _ = ctx.ism_volume(v, dims, isos, extractor, ds, H, isoamd.DF_LBVH,
                                          flags=flags)
        m = st.n_nonempty
        out.append((smap.numpy()[: m * m].reshape(m, m), ids))
    return out


@pytest.mark.parametrize("extractor,H", [(isoamd.EXTRACT_MC, 128), (isoamd.EXTRACT_APPROX, 256),
                                         (isoamd.EXTRACT_MC, 64)])
def test_batch_equals_per_volume(ctx, extractor, H):
    dims_list = [(48, 40, 36), (64, 64, 64), (32, 48, 40), (24, 24, 24)]
    ks = [12, 20, 7, 3]
    imgs = [isoamd.gen_blobs(*d, seed=11 + i) for i, d in enumerate(dims_list)]
    isos_list = [_isos(img, k) for img, k in zip(imgs, ks)]
    vols = [ctx.to_device(img) for img in imgs]
    ds, flags = 2, isoamd.COMPAT_ALL
    print(f"batch: extractor={extractor} H={H} dims={dims_list} k={ks}", flush=True)
    got, st = ctx.ism_batch(vols, dims_list, isos_list, extractor, ds, H, isoamd.DF_LBVH, flags)
    want = _per_volume(ctx, vols, dims_list, isos_list, extractor, ds, H, flags)
    assert st.n_nonempty == sum(w[0].shape[0] for w in want)
    for (gm, gi), (wm, wi) in zip(got, want):
        assert list(gi) == list(wi)
        assert np.array_equal(gm, wm)


def test_batch_drains_16bit_counters(ctx):
    # Planar isosurfaces: every field takes few distinct values, so the
    # diagonal pairs put > 65535 voxels into single bins and the 16-bit LDS
    # counters must drain to the workgroup's global area.
    dims = (48, 192, 192)
    X, Y, Z = dims
    n = X * Y * Z
    # z-fastest image, value = x: planes for the voxel approximation
    img_approx = (np.arange(n, dtype=np.int64) // (Y * Z)).astype(np.float32)
    # compat marching cubes reads x fastest over (X, Y, Z): value = its z
    img_mc = (np.arange(n, dtype=np.int64) // (X * Y)).astype(np.float32)
    isos = np.array([10.5, 20.25, 33.75], np.float32)
    for extractor, img in ((isoamd.EXTRACT_MC, img_mc), (isoamd.EXTRACT_APPROX, img_approx)):
        v = ctx.to_device(img.reshape(dims))
        print(f"batch drain: extractor={extractor} dims={dims}", flush=True)
        got, _ = ctx.ism_batch([v, v], [dims, dims], [isos, isos[:2]], extractor, 1, 256,
                               isoamd.DF_LBVH, isoamd.COMPAT_ALL)
        want = _per_volume(ctx, [v, v], [dims, dims], [isos, isos[:2]], extractor, 1, 256,
                           isoamd.COMPAT_ALL)
        for (gm, gi), (wm, wi) in zip(got, want):
            assert list(gi) == list(wi)
            assert np.array_equal(gm, wm)
            assert np.allclose(np.diag(gm), 1.0, atol=1e-6)


def test_batch_skips_empty_volumes(ctx):
    dims = (24, 20, 16)
    flat = np.zeros(dims, np.float32)
    img = isoamd.gen_blobs(*dims, seed=5)
    vols = [ctx.to_device(flat), ctx.to_device(img)]
    isos = [np.array([0.5, 1.5], np.float32), _isos(img, 5)]
    got, st = ctx.ism_batch(vols, [dims, dims], isos, isoamd.EXTRACT_MC, 1, 32, isoamd.DF_LBVH,
                            isoamd.COMPAT_ALL)
    assert got[0][0].shape == (0, 0) and got[0][1].size == 0
    want = _per_volume(ctx, vols[1:], [dims], isos[1:], isoamd.EXTRACT_MC, 1, 32,
                       isoamd.COMPAT_ALL)
    assert np.array_equal(got[1][0], want[0][0])
    assert st.n_nonempty == want[0][0].shape[0]


@pytest.mark.parametrize("H", [128, 256])
def test_batch_graph_replays_equal_eager(ctx, H):
    # The pair stage captured as a HIP graph on the first call and replayed on
    # later ones, with unrelated eager work (another batch, another layout)
    # between the replays: every replay must give the eager maps bit for bit.
    dims_list = [(40, 36, 32), (56, 48, 40), (24, 32, 28)]
    ks = [9, 14, 5]
    imgs = [isoamd.gen_blobs(*d, seed=51 + i) for i, d in enumerate(dims_list)]
    isos_list = [_isos(img, k) for img, k in zip(imgs, ks)]
    vols = [ctx.to_device(img) for img in imgs]
    ds, flags = 2, isoamd.COMPAT_ALL
    print(f"batch graph: H={H} dims={dims_list} k={ks}", flush=True)
    want, _ = ctx.ism_batch(vols, dims_list, isos_list, isoamd.EXTRACT_MC, ds, H,
                            isoamd.DF_LBVH, flags)
    replays = []
    for rep in range(3):
        got, st = ctx.ism_batch(vols, dims_list, isos_list, isoamd.EXTRACT_MC, ds, H,
                                isoamd.DF_LBVH, flags | isoamd.BATCH_GRAPH)
        replays.append(st.graph_replays)
        for (gm, gi), (wm, wi) in zip(got, want):
            assert list(gi) == list(wi)
            assert np.array_equal(gm, wm), f"replay {rep}"
        # eager work in between: a different batch on the same context
        ctx.ism_batch(vols[:1], dims_list[:1], [isos_list[0][:4]], isoamd.EXTRACT_MC, ds, H,
                      isoamd.DF_LBVH, flags)
        ctx.synchronize()
    assert replays == [0, 1, 2]


def test_batch_graph_layout_with_extra_empty_isovalue(ctx):
    # Same volumes and the same nonempty surfaces, but the last volume gets
    # one more isovalue above its maximum (an empty surface): the task list
    # and field layout stay the same while the marginal bounds move, so the
    # captured pair stage of the first layout must not be replayed for it.
    dims_list = [(40, 36, 32), (32, 40, 28)]
    imgs = [isoamd.gen_blobs(*d, seed=71 + i) for i, d in enumerate(dims_list)]
    isos_list = [_isos(imgs[0], 6), _isos(imgs[1], 5)]
    vols = [ctx.to_device(img) for img in imgs]
    ds, H, flags = 2, 128, isoamd.COMPAT_ALL
    more = [isos_list[0],
            np.append(isos_list[1], np.float32(float(imgs[1].max()) + 1.0)).astype(np.float32)]
    print(f"batch graph layout: dims={dims_list} k=6,5 then 6,6", flush=True)
    for isos in (isos_list, more, isos_list, more):
        want, _ = ctx.ism_batch(vols, dims_list, isos, isoamd.EXTRACT_MC, ds, H,
                                isoamd.DF_LBVH, flags)
        got, _ = ctx.ism_batch(vols, dims_list, isos, isoamd.EXTRACT_MC, ds, H,
                               isoamd.DF_LBVH, flags | isoamd.BATCH_GRAPH)
        for (gm, gi), (wm, wi) in zip(got, want):
            assert list(gi) == list(wi)
            assert np.array_equal(gm, wm)
