"""The 600x248x248 front workload at full size: 100 isovalues, distance fields
on the 150x62x62 grid (ds=4), 256-bin histograms, marching cubes in the
reference's compat mode, whose classify launch covers 33,554,432 of the
36,902,400 cells (isosurfaces_cuda/src/MarchingCube.cpp:599-608).

Against the oracle: vertex sets and cell statistics of two isosurfaces, field
values on sampled voxels, joint counts of sampled pairs, and the whole
100 x 100 map in the reference's float NMI order (JointHistogram.cu:24-60)
over the GPU's fields, bit for bit.  LBVH fields equal brute-force fields on
three isosurfaces."""
import numpy as np
import pytest

import isoamd
import oracle

pytestmark = pytest.mark.gpu

DIMS, K, DS, H = (600, 248, 248), 100, 4, 256
F = (DIMS[0] // DS) * (DIMS[1] // DS) * (DIMS[2] // DS)


@pytest.fixture(scope="module")
def c2(ctx):
    print("c2: generating the front volume and running the ISM", flush=True)
    img = isoamd.gen_front(*DIMS, seed=2)
    isos = np.array([(i + 1) / (K + 1) for i in range(K)], np.float32)
    d = ctx.to_device(img)
    smap, ids, st, fields = ctx.ism_volume(d, DIMS, isos, isoamd.EXTRACT_MC, DS, H,
                                           isoamd.DF_LBVH, isoamd.COMPAT_ALL, want_fields=True)
    m = st.n_nonempty
    out = {"img": img, "isos": isos, "d": d, "m": m, "ids": ids, "st": st,
           "map": smap.numpy()[: m * m].reshape(m, m), "fields": fields.numpy()[:m], "mc": {}}
    smap.free()
    fields.free()
    yield out
    d.free()


def _oracle_mc(c2, i):
    if i not in c2["mc"]:
        X, Y, Z = DIMS
        c2["mc"][i] = oracle.mc(c2["img"].ravel(), DIMS, c2["isos"][i],
                                ncls=oracle.mc_ncls_compat(X, Y, Z))
    return c2["mc"][i]


def test_c2_map_shape_and_diagonal(c2):
    m, mp = c2["m"], c2["map"]
    assert m == K and list(c2["ids"]) == list(range(K))
    assert np.all(np.diag(mp) == 1.0)
    assert np.all((mp >= 0.0) & (mp <= 1.0))
    # (i, j) and (j, i) are separate float chains in the reference's order
    # (measured here up to 5.1e-6 apart): equal within the 1e-5 map tolerance
    np.testing.assert_allclose(mp, mp.T, rtol=1e-5, atol=1e-7)


def test_c2_classify_limit_is_the_reference_launch():
    assert oracle.mc_ncls_compat(*DIMS) == 33554432  # of 36,902,400 cells


@pytest.mark.parametrize("i", [10, 90])
def test_c2_marching_cubes_vertex_sets(ctx, c2, i):
    print(f"c2: marching cubes at isovalue {i} against the oracle", flush=True)
    out, n, stats = ctx.mc_extract(c2["d"], DIMS, float(c2["isos"][i]),
                                   isoamd.COMPAT_MC | isoamd.MC_SORTED)
    want, (active, total) = _oracle_mc(c2, i)
    assert n == want.shape[0] and (stats.active_cells, stats.total_verts) == (active, total)
    assert np.array_equal(out.numpy()[:n], want)
    out.free()


@pytest.mark.parametrize("i", [10, 90])
def test_c2_fields_on_sampled_voxels(c2, i):
    print(f"c2: hierarchy distance field of isovalue {i} on 4096 samples", flush=True)
    pts, _ = _oracle_mc(c2, i)
    rng = np.random.default_rng(i)
    idx = np.sort(rng.choice(F, 4096, replace=False)).astype(np.int64)
    want = oracle.df_tree(oracle.normalize(pts, DIMS), DIMS, DS, idx=idx)
    assert np.array_equal(c2["fields"][i][idx], want)


def test_c2_lbvh_equals_brute_force(ctx, c2):
    sel = np.array([0, 50, 99])
    print("c2: brute-force fields of isovalues 0, 50, 99", flush=True)
    _, bids, bst, bf = ctx.ism_volume(c2["d"], DIMS, c2["isos"][sel], isoamd.EXTRACT_MC, DS, H,
                                      isoamd.DF_BRUTE, isoamd.COMPAT_ALL, want_fields=True)
    assert bst.n_nonempty == 3
    assert np.array_equal(bf.numpy()[:3], c2["fields"][sel])
    bf.free()


@pytest.mark.parametrize("i,j", [(10, 50), (50, 51), (99, 0)])
def test_c2_joint_counts_of_sampled_pairs(ctx, c2, i, j):
    f = c2["fields"]
    d = ctx.to_device(np.stack([f[i], f[j]]))
    _, joint, rows, cols = ctx.hist_nmi(d, 2, F, H, want_joint=True)
    want = oracle.hist_pair(f[i], f[j], H, *oracle.interval(f[i]), *oracle.interval(f[j]))
    assert np.array_equal(joint.numpy()[0, 1], want)
    assert np.array_equal(joint.numpy()[1, 0], want.T)
    d.free()


def test_c2_whole_map_equals_reference_float_order(c2):
    # every ordered pair's histogram and NMI recomputed by the oracle from the
    # GPU's fields, in the reference's float order
    print("c2: oracle map over the 100 GPU fields (10,000 ordered pairs)", flush=True)
    m32, m64 = oracle.ism(c2["fields"], H)
    got = c2["map"]
    diff = np.count_nonzero(got != m32)
    rel = np.abs(m32.astype(np.float64) - m64) / np.abs(m64)
    print(f"c2: entries differing from the reference order: {diff}; the reference order's own "
          f"distance from the exact value: max {rel.max():.3e}, "
          f"{int((rel > 1e-5).sum())} entries above 1e-5", flush=True)
    assert diff == 0
