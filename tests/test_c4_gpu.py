"""The 512^3 workload (sum of sine products, 64 isovalues, DF grid 128^3,
256 bins) at full size: LBVH fields equal brute-force fields bit for bit on
the GPU for two isosurfaces, sampled voxels equal the oracle's hierarchy DF,
and the map has the properties every ISM has."""
import numpy as np
import pytest

import isoamd
import oracle

pytestmark = pytest.mark.gpu


def test_c4_full_size_lbvh_equals_brute_and_oracle(ctx):
    dims, k, ds, H = (512, 512, 512), 64, 4, 256
    X, Y, Z = dims
    F = (X // ds) * (Y // ds) * (Z // ds)
    img = isoamd.gen_sines(*dims, seed=4)
    lo, hi = float(img.min()), float(img.max())
    isos = np.array([lo + (i + 1) * (np.float64(hi) - lo) / (k + 1) for i in range(k)], np.float32)
    d = ctx.to_device(img)
    print("c4: ISM at full size", flush=True)
    smap, ids, st, fields = ctx.ism_volume(d, dims, isos, isoamd.EXTRACT_MC, ds, H,
                                           isoamd.DF_LBVH, isoamd.COMPAT_ALL, want_fields=True)
    m = st.n_nonempty
    assert m == k
    mp = smap.numpy()[: m * m].reshape(m, m)
    assert np.all(np.diag(mp) == 1.0)
    assert np.all((mp >= 0) & (mp <= 1.0))
    lbvh = fields.numpy()[:m]
    # sampled entries, both orientations: the reference's float NMI of the
    # GPU fields (M[i][j] and M[j][i] are separate float chains there)
    rng = np.random.default_rng(6)
    for _ in range(6):
        i, j = (int(q) for q in rng.integers(0, m, 2))
        for a, b in ((i, j), (j, i)):
            h = oracle.hist_pair(lbvh[a], lbvh[b], H, *oracle.interval(lbvh[a]),
                                 *oracle.interval(lbvh[b]))
            assert np.float32(oracle.nmi_ref32(h, F)) == mp[a, b], (a, b)
    # brute force over every sample for two isosurfaces: one small, one large
    sel = np.array([0, k // 2])
    print(f"c4: brute force for isovalues {sel.tolist()}", flush=True)
    _, bids, bst, bfields = ctx.ism_volume(d, dims, isos[sel], isoamd.EXTRACT_MC, ds, H,
                                           isoamd.DF_BRUTE, isoamd.COMPAT_ALL, want_fields=True)
    assert bst.n_nonempty == 2
    assert np.array_equal(bfields.numpy()[:2], lbvh[sel])
    # oracle on sampled voxels of the small isosurface
    print("c4: oracle spot check", flush=True)
    pts, _ = oracle.mc(img.ravel(), dims, isos[0], ncls=oracle.mc_ncls_compat(X, Y, Z))
    rng = np.random.default_rng(0)
    idx = np.sort(rng.choice(F, 4096, replace=False)).astype(np.int64)
    want = oracle.df_tree(oracle.normalize(pts, dims), dims, ds, idx=idx)
    assert np.array_equal(lbvh[0][idx], want)
