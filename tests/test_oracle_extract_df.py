"""Pins the oracle's extraction and distance-field restatements with the
reference's own tables, closed forms and invariants."""
import os
import re

import numpy as np
import pytest

import oracle

REF_TABLES = "/root/reference/isosurfaces_lib/isosurfaces_cuda/src/tables.h"
EDGES = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)]


def test_case_tables_textbook_counts():
    em, nv = oracle.mc_tables()
    assert nv[0] == 0 and nv[255] == 0
    for k in range(8):                       # one corner below iso: one triangle
        assert nv[1 << k] == 3 and bin(int(em[1 << k])).count("1") == 3
    assert nv[0b00000011] == 6               # an edge's two corners: quad = 2 triangles
    assert nv[0b00001111] == 6               # a whole face: plane through 4 edges
    assert nv[0b01000001] == 6               # opposite corners: two separate triangles
    assert nv[0b00000101] == 6               # face diagonal, corners cut off: 2 triangles
    assert nv[0b11111010] == 12              # its complement joins into one hexagon-like piece
    for c in range(256):
        assert nv[c] % 3 == 0 and nv[c] <= 15
        mask = 0
        for e, (a, b) in enumerate(EDGES):
            if ((c >> a) & 1) != ((c >> b) & 1):
                mask |= 1 << e
        assert em[c] == mask


@pytest.mark.skipif(not os.path.exists(REF_TABLES), reason="reference tables not mounted")
def test_case_tables_equal_reference_tables():
    """Known-answer check against the reference's own edgeTable/numVertsTable and
    the edges its triTable uses (isosurfaces_cuda/src/tables.h)."""
    src = open(REF_TABLES).read()

    def arr(name):
        i = src.index(name)
        body = src[src.index("{", i): src.index("};", i)].replace("X", "-1")
        return [int(x, 0) for x in re.findall(r"-?0x[0-9a-fA-F]+|-?\d+", body)]

    edge, tri, nverts = arr("edgeTable["), arr("triTable["), arr("numVertsTable[")
    em, nv = oracle.mc_tables()
    assert list(em) == edge and list(nv) == nverts
    for c in range(256):
        used = 0
        for v in tri[16 * c: 16 * c + 16]:
            if v < 0:
                break
            used |= 1 << v
        assert used == em[c]  # every crossed edge yields a vertex


def _ramp(gx, gy, gz):
    # value = logical x in the marching-cubes reading frame (x fastest)
    return np.broadcast_to(np.arange(gx, dtype=np.float32)[None, None, :], (gz, gy, gx)).copy()


def test_mc_ramp_closed_form_with_border_cells():
    gx, gy, gz = 9, 6, 5
    vol = _ramp(gx, gy, gz)
    v, (active, total) = oracle.mc(vol.ravel(), (gx, gy, gz), 3.5)
    # x-edges from 3 to 4 at every (y, z) including the clamped border rows y=gy, z=gz
    assert v.shape[0] == (gy + 1) * (gz + 1)
    assert np.all(v[:, 0] == 3.5)
    assert set(map(tuple, v[:, 1:].astype(int))) == {(y, z) for y in range(gy + 1) for z in range(gz + 1)}
    assert active == gy * gz and total == 6 * gy * gz  # each cell: one quad


def test_mc_grid_point_vertices_deduplicated():
    # iso equal to sample values: vertices land on grid points (t = 1)
    gx, gy, gz = 6, 5, 4
    vol = _ramp(gx, gy, gz)
    v, _ = oracle.mc(vol.ravel(), (gx, gy, gz), 3.0)
    assert np.all(v[:, 0] == 3.0)
    assert len({tuple(r) for r in v}) == v.shape[0] == (gy + 1) * (gz + 1)


def test_mc_classify_limit():
    # 17*9*5 = 765 cells, the reference launch covers 5*128 = 640
    assert oracle.mc_ncls_compat(17, 9, 5) == 640
    assert oracle.mc_ncls_compat(64, 64, 64) == 64 ** 3
    assert oracle.mc_ncls_compat(600, 248, 248) == 33554432
    assert oracle.mc_ncls_compat(300, 200, 150) == 128 * 32768 * 2


def test_mc_vertices_on_edges_interpolate_iso():
    rng = np.random.default_rng(0)
    gx, gy, gz = 12, 10, 8
    vol = rng.random((gz, gy, gx)).astype(np.float32)
    iso = 0.5
    v, (active, total) = oracle.mc(vol.ravel(), (gx, gy, gz), iso)
    assert total % 3 == 0 and active > 0 and v.shape[0] > 0
    sample = lambda x, y, z: vol[min(z, gz - 1), min(y, gy - 1), min(x, gx - 1)]
    for p in v:
        ints = [float(c).is_integer() for c in p]
        assert sum(ints) >= 2
        if sum(ints) == 3:
            continue
        a = ints.index(False)
        lo = [int(c) for c in p]
        hi = list(lo)
        hi[a] += 1
        f0, f1 = sample(*lo), sample(*hi)
        assert (f0 < iso) != (f1 < iso)
        t = p[a] - lo[a]
        assert abs(f0 + t * (f1 - f0) - iso) < 1e-5


def test_mc_sphere_vertices_near_sphere():
    n = 24
    g = np.arange(n, dtype=np.float64)
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    vol = np.sqrt((x - 11.3) ** 2 + (y - 12.1) ** 2 + (z - 10.7) ** 2).astype(np.float32)
    v, _ = oracle.mc(vol.ravel(), (n, n, n), 7.0)
    r = np.sqrt(((v - [11.3, 12.1, 10.7]) ** 2).sum(1))
    assert v.shape[0] > 500 and np.all(np.abs(r - 7.0) < 0.08)


def test_approx_closed_forms():
    X, Y, Z = 7, 5, 4
    img = np.broadcast_to(np.arange(X, dtype=np.float32)[:, None, None], (X, Y, Z)).copy()
    c = oracle.approx(img, 2.5)
    assert c.shape[0] == (Y - 1) * (Z - 1) and np.all(c[:, 0] == 2)
    # x-major / z-fastest order
    assert [tuple(r) for r in c] == [(2, y, z) for y in range(Y - 1) for z in range(Z - 1)]
    # corners equal to iso count as neither above nor below
    assert oracle.approx(img, 2.0).shape[0] == 0


def test_approx_single_cell_all_patterns():
    for pattern in range(256):
        vals = np.array([(pattern >> k) & 1 for k in range(8)], np.float32)
        img = vals.reshape(2, 2, 2)
        got = oracle.approx(img, 0.5).shape[0]
        assert got == (1 if 0 < pattern < 255 else 0)
        img2 = img.copy()
        img2[0, 0, 0] = 0.5  # a corner at iso never decides
        want = int(any(img2.ravel() > 0.5) and any(img2.ravel() < 0.5))
        assert oracle.approx(img2, 0.5).shape[0] == want


def test_normalisation_and_sample_points():
    pts = np.array([[3, 4, 5], [0, 0, 0]], np.int32)
    out = oracle.normalize(pts, (10, 20, 8))
    assert np.array_equal(out, np.array([[3, 4, 5], [0, 0, 0]], np.float32) / np.float32(20))
    s = oracle.sample_point(0, (10, 20, 8), 4)
    assert np.array_equal(s, np.array([2 / 10, 2 / 20, 2 / 8], np.float32))
    idx = 1 * (5 * 2) + 2 * 2 + 1  # x=1, y=2, z=1 on the (2,5,2) grid
    s = oracle.sample_point(idx, (10, 20, 8), 4)
    assert np.array_equal(s, np.array([6 / 10, 10 / 20, 6 / 8], np.float32))


def test_df_plane_closed_form():
    n = 16
    yy, zz = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    pts = np.stack([np.full(n * n, 5), yy.ravel(), zz.ravel()], 1).astype(np.int32)
    df = oracle.df_brute(oracle.normalize(pts, (n, n, n)), (n, n, n), 1).reshape(n, n, n)
    sx = (np.arange(n, dtype=np.float32) / np.float32(n))
    want = np.abs(sx - np.float32(5) / np.float32(n))
    assert np.array_equal(df, np.broadcast_to(want[:, None, None], (n, n, n)))


def test_df_matches_float64_euclid_and_single_point_quirk():
    rng = np.random.default_rng(1)
    dims = (20, 14, 10)
    pts = (rng.random((37, 3)) * [20, 14, 10]).astype(np.float32)
    pn = oracle.normalize(pts, dims)
    df = oracle.df_brute(pn, dims, 2)
    idx = np.arange(df.size)
    S = np.stack([oracle.sample_point(i, dims, 2) for i in idx]).astype(np.float64)
    ref = np.sqrt(((S[:, None, :] - pn[None].astype(np.float64)) ** 2).sum(-1)).min(1)
    np.testing.assert_allclose(df, ref, rtol=2e-6, atol=1e-7)
    tree = oracle.df_tree(pn, dims, 2)          # hierarchy returns the same exact minimum
    assert np.array_equal(tree, df)
    # the reference's pre-sample bound first (BVH.cu:640-644): the same minimum
    assert np.array_equal(oracle.df_tree(pn, dims, 2, presample=True), df)
    big = (rng.random((5000, 3)) * [20, 14, 10]).astype(np.float32)
    bn = oracle.normalize(big, dims)
    assert np.array_equal(oracle.df_tree(bn, dims, 2, presample=True), oracle.df_brute(bn, dims, 2))
    one = oracle.df_brute(pn[:1], dims, 2)   # P == 1 keeps d^2 (BVH.cu:604-611)
    ref1 = ((S - pn[0].astype(np.float64)) ** 2).sum(1)
    np.testing.assert_allclose(one, ref1, rtol=2e-6, atol=1e-7)
