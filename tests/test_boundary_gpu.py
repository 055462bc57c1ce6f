"""The two remaining entry points of the reference's CUDA library that its
host code links against, on the GPU against the oracle, bit for bit:
calculate_distances_h (brute force in voxel units, 2M-point runs;
isosurfaces_cuda/src/DistanceField.cu:15-98) and findExtrema_h (values of
local extrema; src/IsosurfaceExtrema.cu:22-138), through the C ABI and
through the reference-named mirror."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import isoamd
import oracle
from conftest import PKG

pytestmark = pytest.mark.gpu
DRIVER = os.path.join(PKG, "host", "compat_driver")


@pytest.mark.parametrize("grid,ds,n", [((20, 16, 12), 2, 5000), ((9, 7, 5), 3, 1),
                                       ((33, 10, 4), 1, 777)])
def test_df_voxel_units_matches_oracle(ctx, grid, ds, n):
    rng = np.random.default_rng(n)
    pts = (rng.random((n, 3)) * np.array(grid) * ds).astype(np.float32)
    d = ctx.to_device(pts)
    got = ctx.df_voxel_units(d, n, grid, ds).numpy()
    assert np.array_equal(got, oracle.df_voxel_units(pts, grid, ds))


def test_df_voxel_units_point_runs_and_empty_set(ctx):
    # 2,100,000 points: two runs of the reference's launch loop, the second
    # starting from the square of the first's root
    grid, ds, n = (8, 6, 5), 4, 2_100_000
    rng = np.random.default_rng(5)
    pts = (rng.random((n, 3)) * 40.0 - 4.0).astype(np.float32)
    d = ctx.to_device(pts)
    got = ctx.df_voxel_units(d, n, grid, ds).numpy()
    assert np.array_equal(got, oracle.df_voxel_units(pts, grid, ds))
    empty = ctx.df_voxel_units(None, 0, grid, ds).numpy()
    assert np.all(empty == np.sqrt(np.float32(1e38)))


@pytest.mark.parametrize("dims", [(24, 24, 24), (30, 20, 12), (7, 5, 3)])
def test_find_extrema_matches_oracle(ctx, dims):
    X, Y, Z = dims
    assert (X - 1) + (Y - 1) * Y + (Z - 1) * Y * Z < X * Y * Z
    rng = np.random.default_rng(X)
    # coarse values: ties exercise the "more than 5 equal" rule
    vol = np.round(rng.random(X * Y * Z) * 6).astype(np.float32)
    vol[::97] = -2e10  # values at or below -1e10 never come back
    got = ctx.find_extrema(ctx.to_device(vol), dims)
    want = oracle.find_extrema(vol, dims)
    assert got.size > 0 and np.array_equal(got, want)


def test_find_extrema_rejects_dims_the_reference_reads_past(ctx):
    vol = np.zeros(3 * 3 * 6, np.float32)
    with pytest.raises(isoamd.IsoError) as e:
        ctx.find_extrema(ctx.to_device(vol), (3, 3, 6))
    assert e.value.status == isoamd.ERR_UNSUPPORTED


class _Int3(C.Structure):
    _fields_ = [("x", C.c_int), ("y", C.c_int), ("z", C.c_int)]


def test_find_extrema_h_mirror():
    lib = C.CDLL(os.path.join(PKG, "libisosurfaces_compat.so"))
    lib.findExtrema_h.restype = C.c_long
    lib.findExtrema_h.argtypes = [C.c_void_p, _Int3, C.POINTER(C.POINTER(C.c_float))]
    dims = (16, 16, 16)
    vol = isoamd.gen_blobs(*dims, seed=3)
    vol = np.round(vol.ravel() * 20).astype(np.float32)
    out = C.POINTER(C.c_float)()
    n = lib.findExtrema_h(vol.ctypes.data, _Int3(*dims), C.byref(out))
    got = np.ctypeslib.as_array(out, shape=(n,)).copy() if n else np.zeros(0, np.float32)
    C.CDLL(None).free(out)  # the caller frees, as with approximate_isosurfaces_h
    assert np.array_equal(got, oracle.find_extrema(vol, dims))


def test_calculate_distances_h_mirror(tmp_path):
    grid, ds, n = (12, 10, 8), 2, 3000
    rng = np.random.default_rng(8)
    pts = (rng.random((n, 3)) * 20.0).astype(np.float32)
    src, out = tmp_path / "pts.f32", tmp_path / "df.f32"
    pts.tofile(src)
    r = subprocess.run([DRIVER, "dist", str(src), str(n), *map(str, grid), str(ds), str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert np.array_equal(np.fromfile(out, np.float32), oracle.df_voxel_units(pts, grid, ds))
