"""The reference-interface mirror (libisosurfaces_compat.so: approximate_isosurfaces_h,
calculate_histogram_h, generateHierarchy_h/_W/i3_W, cudaMarchingCube) against the
oracle, through the same names and ownership rules the reference's callers use."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import isoamd
import oracle
from conftest import PKG

pytestmark = pytest.mark.gpu
COMPAT = os.path.join(PKG, "libisosurfaces_compat.so")
DRIVER = os.path.join(PKG, "host", "compat_driver")


@pytest.fixture(scope="module")
def compat():
    L = C.CDLL(COMPAT)
    L.approximate_isosurfaces_h.restype = C.c_long
    L.approximate_isosurfaces_h.argtypes = [C.c_void_p, C.c_float, C.c_int, C.c_int, C.c_int, C.c_int,
                                            C.POINTER(C.c_void_p)]
    L.calculate_histogram_h.restype = None
    L.calculate_histogram_h.argtypes = [C.POINTER(C.c_void_p), C.c_void_p, C.c_void_p, C.c_int,
                                        C.c_ulong, C.c_ulong, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_int, C.c_bool]
    return L


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def test_approximate_isosurfaces_h(compat):
    dims = (33, 27, 21)
    img = isoamd.gen_blobs(*dims, seed=12)
    iso = np.float32(img.mean())
    out = C.c_void_p()
    n = compat.approximate_isosurfaces_h(_p(img), iso, *dims, 1, C.byref(out))
    want = oracle.approx(img, iso)
    got = np.ctypeslib.as_array(C.cast(out, C.POINTER(C.c_int32)), shape=(n, 3)).copy()
    C.CDLL(None).free(out)  # caller frees, as Single.cpp:227-230 does
    assert n == want.shape[0] and np.array_equal(got, want)


@pytest.mark.parametrize("multi", [False, True])
def test_calculate_histogram_h(compat, multi):
    n, F, H = 5, 30000, 128
    rng = np.random.default_rng(4)
    nf = 2 * n if multi else n
    fields = np.abs(rng.random((nf, F)) - rng.random((nf, 1))).astype(np.float32)
    iv = np.array([oracle.interval(f) for f in fields], np.float32)
    mins, maxs = np.ascontiguousarray(iv[:, 0]), np.ascontiguousarray(iv[:, 1])
    ptrs = (C.c_void_p * nf)(*[f.ctypes.data for f in fields])
    simmap = np.full((n, n), -1, np.float32)
    cols = np.zeros(n * n * H, np.uint32)
    rows = np.zeros(n * n * H, np.uint32)
    joint = np.full(4, 7, np.uint32)  # never written by the reference
    compat.calculate_histogram_h(ptrs, _p(mins), _p(maxs), F, n, H, _p(joint), _p(cols), _p(rows),
                                 _p(simmap), n, multi)
    m32, _ = oracle.ism(fields, H, multi=multi)  # the reference's float order
    assert np.array_equal(simmap, m32)
    assert np.all(joint == 7)
    for i in range(n):
        for j in range(n):
            b = n + j if multi else j
            h = oracle.hist_pair(fields[i], fields[b], H, *iv[i], *iv[b])
            o = (i * n + j) * H
            assert np.array_equal(rows[o:o + H], h.sum(1)) and np.array_equal(cols[o:o + H], h.sum(0))


def _run(args, tmp_path):
    r = subprocess.run([DRIVER] + [str(a) for a in args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return int(r.stdout.strip())


def test_cuda_marching_cube_mirror(tmp_path):
    dims = (30, 22, 18)
    img = isoamd.gen_blobs(*dims, seed=13)
    iso = np.float32(img.min() + 0.45 * (img.max() - img.min()))
    vol = tmp_path / "vol.f32"
    img.tofile(vol)
    out = tmp_path / "out.f32"
    n = _run(["mc", vol, *dims, repr(float(iso)), out], tmp_path)
    got = np.fromfile(out, np.float32).reshape(n, 4)
    want, _ = oracle.mc(img.ravel(), dims, iso, ncls=oracle.mc_ncls_compat(*dims))
    assert np.array_equal(got[:, :3], want) and np.all(got[:, 3] == 1.0)


@pytest.mark.parametrize("kind", ["df", "dfh", "dfi3"])
def test_generate_hierarchy_mirror(tmp_path, kind):
    dims, ds = (40, 30, 20), 2
    img = isoamd.gen_blobs(*dims, seed=14)
    iso = np.float32(img.mean())
    if kind == "dfi3":
        pts = oracle.approx(img, iso)
    else:
        pts, _ = oracle.mc(img.ravel(), dims, iso, ncls=oracle.mc_ncls_compat(*dims))
    norm = oracle.normalize(pts, dims)
    src = tmp_path / "pts.bin"
    (norm if kind == "dfh" else pts).tofile(src)
    out = tmp_path / "df.f32"
    _run([kind, src, pts.shape[0], *dims, ds, out], tmp_path)
    got = np.fromfile(out, np.float32)
    assert np.array_equal(got, oracle.df_brute(norm, dims, ds))
