"""HIP marching cubes and voxel approximation against the oracle: the vertex
set (sorted), the cell list and the per-case counts must be bit-exact."""
import numpy as np
import pytest

import isoamd
import oracle

pytestmark = pytest.mark.gpu


def _blobs(X, Y, Z, seed):
    return isoamd.gen_blobs(X, Y, Z, nblobs=8, noise=0.01, seed=seed)


def _isos(img, k):
    lo, hi = float(img.min()), float(img.max())
    return [np.float32(lo + (i + 1) * (hi - lo) / (k + 1)) for i in range(k)]


CASES = [((64, 64, 64), 1, 4), ((17, 9, 5), 2, 3), ((40, 33, 20), 3, 3), ((31, 64, 23), 4, 2)]


@pytest.mark.parametrize("dims,seed,k", CASES)
@pytest.mark.parametrize("compat", [True, False])
def test_mc_matches_oracle(ctx, dims, seed, k, compat):
    X, Y, Z = dims
    print(f"mc parity dims={dims} seed={seed} compat={compat}", flush=True)
    img = _blobs(X, Y, Z, seed)
    d = ctx.to_device(img)
    flags = (isoamd.COMPAT_MC if compat else 0) | isoamd.MC_SORTED
    for iso in _isos(img, k):
        out, n, st = ctx.mc_extract(d, (X, Y, Z), float(iso), flags)
        got = out.numpy()[:n] if n else np.zeros((0, 3), np.float32)
        if compat:
            # the reference reads the z-fastest image x-fastest with dims (X,Y,Z)
            want, (active, total) = oracle.mc(img.ravel(), (X, Y, Z), iso,
                                              ncls=oracle.mc_ncls_compat(X, Y, Z))
        else:
            # the image read in its own (x,y,z) order: transpose to x-fastest,
            # interior cells only (no clamped border)
            want, (active, total) = _oracle_fixed(img, iso)
        assert n == want.shape[0]
        assert np.array_equal(got, want)
        assert st.active_cells == active and st.total_verts == total
        assert st.triangles * 3 == st.total_verts


def _oracle_fixed(img, iso):
    """The image read in its own (x,y,z) frame, interior cells only."""
    X, Y, Z = img.shape
    xf = np.ascontiguousarray(img.transpose(2, 1, 0))  # x fastest
    return oracle.mc(xf.ravel(), (X, Y, Z), iso, interior=True)


def test_mc_classify_limit_truncates(ctx):
    X, Y, Z = 300, 200, 150  # > 65535 classify blocks: the reference skips the tail
    img = isoamd.gen_sines(X, Y, Z, seed=11)
    d = ctx.to_device(img)
    iso = 0.3
    out, n, st = ctx.mc_extract(d, (X, Y, Z), iso, isoamd.COMPAT_MC | isoamd.MC_SORTED)
    want, (active, total) = oracle.mc(img.ravel(), (X, Y, Z), np.float32(iso),
                                      ncls=oracle.mc_ncls_compat(X, Y, Z))
    assert n == want.shape[0] and np.array_equal(out.numpy()[:n], want)
    assert st.active_cells == active and st.total_verts == total


def test_mc_grid_point_vertices(ctx):
    # isovalue equal to many samples: vertices on grid points, deduplicated across edges
    X, Y, Z = 24, 20, 16
    rng = np.random.default_rng(3)
    img = rng.integers(0, 4, size=(X, Y, Z)).astype(np.float32)
    d = ctx.to_device(img)
    for iso in (1.0, 2.0, 2.5):
        out, n, st = ctx.mc_extract(d, (X, Y, Z), iso, isoamd.COMPAT_MC | isoamd.MC_SORTED)
        want, (active, total) = oracle.mc(img.ravel(), (X, Y, Z), np.float32(iso),
                                          ncls=oracle.mc_ncls_compat(X, Y, Z))
        assert n == want.shape[0] and np.array_equal(out.numpy()[:n], want)
        assert st.total_verts == total


def test_mc_empty_and_capacity(ctx):
    img = _blobs(20, 20, 20, 5)
    d = ctx.to_device(img)
    out, n, st = ctx.mc_extract(d, (20, 20, 20), float(img.max()) + 1.0)
    assert n == 0 and st.total_verts == 0
    small = ctx.empty((4, 3), np.float32)
    with pytest.raises(isoamd.IsoError) as e:
        isoamd.lib()  # capacity error surfaces as status 4 when the wrapper cannot grow
        import ctypes as C
        cnt = C.c_int64()
        st2 = isoamd.McStats()
        s = isoamd.lib().iso_mc_extract(ctx.h, C.c_void_p(d.ptr), 20, 20, 20, C.c_float(float(img.mean())),
                                        isoamd.COMPAT_MC, C.c_void_p(small.ptr), 4, C.byref(cnt), C.byref(st2))
        ctx._check(s, "capacity")
    assert e.value.status == isoamd.ERR_CAPACITY and cnt.value > 4


@pytest.mark.parametrize("dims,seed", [((64, 64, 64), 1), ((17, 9, 5), 2), ((50, 31, 26), 6)])
def test_approx_matches_oracle(ctx, dims, seed):
    X, Y, Z = dims
    img = _blobs(X, Y, Z, seed)
    d = ctx.to_device(img)
    for iso in _isos(img, 3):
        out, n = ctx.approx_extract(d, dims, float(iso))
        want = oracle.approx(img, iso)
        assert n == want.shape[0]
        assert np.array_equal(out.numpy()[:n], want)


def test_index_limits_are_errors(ctx):
    # Grids whose 32-bit candidate vertex ids (marching cubes: grid point * 5
    # + edge kind) or cell offsets (voxel approximation) would wrap are
    # refused before any kernel runs, so the small buffer is never read.
    d = ctx.to_device(np.zeros((4, 4, 4), np.float32))
    print("limits: marching cubes 1000^3, approximation 2000x2000x1100", flush=True)
    with pytest.raises(isoamd.IsoError) as e:
        ctx.mc_extract(d, (1000, 1000, 1000), 0.5)
    assert e.value.status == isoamd.ERR_UNSUPPORTED
    with pytest.raises(isoamd.IsoError) as e:
        ctx.approx_extract(d, (2000, 2000, 1100), 0.5)
    assert e.value.status == isoamd.ERR_UNSUPPORTED
    # the context stays usable
    img = isoamd.gen_blobs(20, 18, 16, seed=3)
    out, n, st = ctx.mc_extract(ctx.to_device(img), (20, 18, 16), float(img.mean()))
    assert n > 0
