"""ctypes binding of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, always as the checker.  The product library never uses it.  See
iso_oracle.h for what each function restates from the reference.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        vp, i32, i64, f32 = C.c_void_p, C.c_int, C.c_int64, C.c_float
        sig = {
            "orc_ingest": (None, [vp, i32, i32, i32, f32, f32, vp, vp, vp]),
            "orc_equalized_isovalues": (None, [vp, i64, f32, f32, i32, vp]),
            "orc_mc_tables": (None, [vp, vp]),
            "orc_mc_ncls_compat": (i64, [i32, i32, i32]),
            "orc_mc": (i64, [vp, i32, i32, i32, f32, i64, i32, vp, i64, vp]),
            "orc_approx": (i64, [vp, i32, i32, i32, f32, vp, i64]),
            "orc_normalize_f3": (None, [vp, i64, i32, i32, i32, vp]),
            "orc_normalize_i3": (None, [vp, i64, i32, i32, i32, vp]),
            "orc_sample_point": (None, [i64, i32, i32, i32, i32, vp]),
            "orc_d2": (f32, [vp, vp]),
            "orc_df_brute": (None, [vp, i64, i32, i32, i32, i32, vp, i32]),
            "orc_df_brute_subset": (None, [vp, i64, i32, i32, i32, i32, vp, i64, vp, i32]),
            "orc_df_tree": (None, [vp, i64, i32, i32, i32, i32, vp, i64, vp, i32]),
            "orc_df_tree_ref": (None, [vp, i64, i32, i32, i32, i32, vp, i64, vp, i32]),
            "orc_interval": (None, [vp, i64, vp, vp]),
            "orc_df_voxel_units": (None, [vp, i64, i32, i32, i32, i32, vp, i32]),
            "orc_find_extrema": (i64, [vp, i32, i32, i32, vp, i64]),
            "orc_hist_pair": (None, [vp, vp, i64, i32, f32, f32, f32, f32, vp]),
            "orc_hist_processed": (i64, [i64]),
            "orc_nmi_ref32": (f32, [vp, i32, i64]),
            "orc_nmi_ref32_ulp": (f32, [vp, i32, i64, C.c_uint64]),
            "orc_logf_cr_margin": (i64, [C.c_uint32, C.c_uint32, C.c_double, i32,
                                         C.POINTER(C.c_uint32)]),
            "orc_nmi_f64": (C.c_double, [vp, i32, i64]),
            "orc_ism": (None, [vp, i32, i64, i32, i32, i32, vp, vp, i32]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(_lib, name)
            fn.restype = res
            fn.argtypes = args
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def nthreads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


def ingest(file_xfast: np.ndarray, dims, cfg_min: float, cfg_max: float):
    """(z-fastest image (X,Y,Z), real_min, real_max) from an x-fastest buffer."""
    X, Y, Z = dims
    f = np.ascontiguousarray(file_xfast, np.float32).ravel()
    img = np.empty((X, Y, Z), np.float32)
    lo = np.zeros(1, np.float32)
    hi = np.zeros(1, np.float32)
    lib().orc_ingest(_p(f), X, Y, Z, cfg_min, cfg_max, _p(img), _p(lo), _p(hi))
    return img, float(lo[0]), float(hi[0])


def equalized_isovalues(img: np.ndarray, real_min: float, real_max: float, k: int) -> np.ndarray:
    d = np.ascontiguousarray(img, np.float32).ravel()
    out = np.empty(k, np.float32)
    lib().orc_equalized_isovalues(_p(d), d.size, real_min, real_max, k, _p(out))
    return out


def mc_tables():
    em = np.zeros(256, np.uint16)
    nv = np.zeros(256, np.uint8)
    lib().orc_mc_tables(_p(em), _p(nv))
    return em, nv


def mc_ncls_compat(gx, gy, gz) -> int:
    return int(lib().orc_mc_ncls_compat(gx, gy, gz))


def mc(vol: np.ndarray, dims, iso: float, ncls: int | None = None, interior: bool = False):
    """Sorted unique vertices (V x 3 float32) and (active_cells, total_verts).
    vol is read x-fastest with dims (gx, gy, gz)."""
    gx, gy, gz = dims
    vol = np.ascontiguousarray(vol, np.float32)
    if ncls is None:
        ncls = gx * gy * gz
    stats = np.zeros(2, np.int64)
    n = lib().orc_mc(_p(vol), gx, gy, gz, iso, ncls, int(interior), None, 0, _p(stats))
    if n < 0:
        raise MemoryError("orc_mc")
    out = np.empty((max(n, 1), 3), np.float32)
    n = lib().orc_mc(_p(vol), gx, gy, gz, iso, ncls, int(interior), _p(out), n, _p(stats))
    return out[:n], (int(stats[0]), int(stats[1]))


def approx(img: np.ndarray, iso: float) -> np.ndarray:
    X, Y, Z = img.shape
    img = np.ascontiguousarray(img, np.float32)
    n = lib().orc_approx(_p(img), X, Y, Z, iso, None, 0)
    out = np.empty((max(n, 1), 3), np.int32)
    lib().orc_approx(_p(img), X, Y, Z, iso, _p(out), n)
    return out[:n]


def normalize(pts: np.ndarray, dims) -> np.ndarray:
    X, Y, Z = dims
    pts = np.ascontiguousarray(pts)
    out = np.empty(pts.shape, np.float32)
    if pts.dtype == np.int32:
        lib().orc_normalize_i3(_p(pts), pts.shape[0], X, Y, Z, _p(out))
    else:
        pts = pts.astype(np.float32)
        lib().orc_normalize_f3(_p(pts), pts.shape[0], X, Y, Z, _p(out))
    return out


def sample_point(idx: int, dims, ds: int) -> np.ndarray:
    out = np.empty(3, np.float32)
    lib().orc_sample_point(idx, *dims, ds, _p(out))
    return out


def df_brute(pts_norm: np.ndarray, dims, ds: int) -> np.ndarray:
    X, Y, Z = dims
    F = (X // ds) * (Y // ds) * (Z // ds)
    out = np.empty(F, np.float32)
    p = np.ascontiguousarray(pts_norm, np.float32)
    lib().orc_df_brute(_p(p), p.shape[0], X, Y, Z, ds, _p(out), nthreads())
    return out


def df_brute_subset(pts_norm: np.ndarray, dims, ds: int, idx: np.ndarray) -> np.ndarray:
    X, Y, Z = dims
    idx = np.ascontiguousarray(idx, np.int64)
    out = np.empty(idx.shape[0], np.float32)
    p = np.ascontiguousarray(pts_norm, np.float32)
    lib().orc_df_brute_subset(_p(p), p.shape[0], X, Y, Z, ds, _p(idx), idx.shape[0], _p(out),
                              nthreads())
    return out


def df_tree(pts_norm: np.ndarray, dims, ds: int, idx: np.ndarray | None = None,
            threads: int | None = None, presample: bool = False) -> np.ndarray:
    """Distance field through the CPU hierarchy (same exact minimum);
    presample: seeded by the reference's 1500-point pre-sample bound first."""
    X, Y, Z = dims
    F = (X // ds) * (Y // ds) * (Z // ds)
    p = np.ascontiguousarray(pts_norm, np.float32)
    fn = lib().orc_df_tree_ref if presample else lib().orc_df_tree
    if idx is None:
        out = np.empty(F, np.float32)
        fn(_p(p), p.shape[0], X, Y, Z, ds, None, 0, _p(out), threads or nthreads())
    else:
        idx = np.ascontiguousarray(idx, np.int64)
        out = np.empty(idx.shape[0], np.float32)
        fn(_p(p), p.shape[0], X, Y, Z, ds, _p(idx), idx.shape[0], _p(out), threads or nthreads())
    return out


def df_voxel_units(pts: np.ndarray, grid, ds: int) -> np.ndarray:
    """calculate_distances_h semantics (voxel units, 2M-point runs)."""
    nx, ny, nz = grid
    p = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
    out = np.empty(nx * ny * nz, np.float32)
    lib().orc_df_voxel_units(_p(p), p.shape[0], nx, ny, nz, ds, _p(out), nthreads())
    return out


def find_extrema(data: np.ndarray, dims) -> np.ndarray:
    """findExtrema_h semantics: extremum values in voxel order."""
    X, Y, Z = dims
    d = np.ascontiguousarray(data, np.float32).ravel()
    n = int(lib().orc_find_extrema(_p(d), X, Y, Z, None, 0))
    out = np.empty(max(n, 1), np.float32)
    lib().orc_find_extrema(_p(d), X, Y, Z, _p(out), n)
    return out[:n]


def interval(f: np.ndarray):
    f = np.ascontiguousarray(f, np.float32)
    lo = np.zeros(1, np.float32)
    hi = np.zeros(1, np.float32)
    lib().orc_interval(_p(f), f.size, _p(lo), _p(hi))
    return float(lo[0]), float(hi[0])


def hist_pair(f1, f2, H, mn1, mx1, mn2, mx2, processed=None) -> np.ndarray:
    f1 = np.ascontiguousarray(f1, np.float32)
    f2 = np.ascontiguousarray(f2, np.float32)
    if processed is None:
        processed = f1.size
    h = np.zeros((H, H), np.uint32)
    lib().orc_hist_pair(_p(f1), _p(f2), processed, H, mn1, mx1, mn2, mx2, _p(h))
    return h


def hist_processed(F: int) -> int:
    return int(lib().orc_hist_processed(F))


def nmi_ref32(hist: np.ndarray, N: int) -> float:
    h = np.ascontiguousarray(hist, np.uint32)
    return float(lib().orc_nmi_ref32(_p(h), h.shape[0], N))


def nmi_ref32_ulp(hist: np.ndarray, N: int, seed: int) -> float:
    """Reference-order NMI with every logf moved by -1/0/+1 ulp (fixed per
    argument and seed): another logf within CUDA's documented 1 ulp."""
    h = np.ascontiguousarray(hist, np.uint32)
    return float(lib().orc_nmi_ref32_ulp(_p(h), h.shape[0], N, seed))


def logf_cr_margin(lo: float, hi: float, rel: float = 2.0 ** -48, integers_only: bool = True):
    """(count, first failing float) of the floats in [lo, hi] whose logf is
    not pinned to one correctly rounded value by any double logarithm."""
    lo_b = int(np.float32(lo).view(np.uint32))
    hi_b = int(np.float32(hi).view(np.uint32))
    first = C.c_uint32(0)
    n = int(lib().orc_logf_cr_margin(lo_b, hi_b, rel, int(integers_only), C.byref(first)))
    return n, (float(np.uint32(first.value).view(np.float32)) if n else None)


def nmi_f64(hist: np.ndarray, N: int) -> float:
    h = np.ascontiguousarray(hist, np.uint32)
    return float(lib().orc_nmi_f64(_p(h), h.shape[0], N))


def ism(fields: np.ndarray, H: int, multi: bool = False, compat_chunks: bool = True):
    """(map_ref32, map_f64) for fields of shape (n or 2n, F)."""
    fields = np.ascontiguousarray(fields, np.float32)
    nf, F = fields.shape
    n = nf // 2 if multi else nf
    m32 = np.zeros((n, n), np.float32)
    m64 = np.zeros((n, n), np.float64)
    lib().orc_ism(_p(fields), n, F, H, int(multi), int(compat_chunks), _p(m32), _p(m64), nthreads())
    return m32, m64


def representatives(smap: np.ndarray, cnt: int, ids=None):
    """Bruckner-Moeller selection as types/src/SimilarityMap.cpp:86-199 states
    it (pure Python over float32 scalars; small maps only).

    prioritize (:86-158): over a set S (in its order), the mean similarity of
    each member to S is a float32 sum in S's order divided by float(|S|); the
    first strict maximum (start -1.0) is pushed with priority float(|S|) * mean,
    then S splits into the members with a smaller map index and the rest.
    selectValues (:162-199): cnt times take the maximum priority and divide
    every other priority by (1 + similarity to the taken one).  The reference's
    heap order for equal priorities is not pinned; here the lower map index
    wins.  Returns [(id, map index, priority)] in the order taken."""
    f32 = np.float32
    smap = np.asarray(smap, np.float32)
    m = smap.shape[0]
    ids = list(range(m)) if ids is None else [int(v) for v in ids]
    heap = []

    def prioritize(elems):
        if not elems:
            return
        n = len(elems)
        best, best_avg = -1, f32(-1.0)
        for a in elems:
            s = f32(0.0)
            for b in elems:
                s = f32(s + smap[a, b])
            avg = f32(s / f32(n))
            if avg > best_avg:
                best, best_avg = a, avg
        heap.append([f32(f32(n) * best_avg), ids[best], best])
        prioritize([e for e in elems if e < best])
        prioritize([e for e in elems if e > best])

    prioritize(list(range(m)))
    out = []
    for _ in range(min(cnt, m)):
        top = max(range(len(heap)), key=lambda i: (heap[i][0], -heap[i][2]))
        p, ident, row = heap.pop(top)
        out.append((ident, row, float(p)))
        for h in heap:
            h[0] = f32(h[0] / f32(f32(1.0) + smap[row, h[2]]))
    return out
