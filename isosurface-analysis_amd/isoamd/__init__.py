"""Python binding of libiso_amd.so (the C ABI in include/iso_amd.h).

A thin ctypes layer used by the tests and bench.py: it owns no compute, every
call goes to the HIP library.  Loading fails loudly if the library has not been
built (``python -c "import __graft_entry__; __graft_entry__.build()"``).

Reference names this layer mirrors (isosurfaces_lib/ of the reference):
  mc_extract       <- cudaMarchingCube          (isosurfaces_cuda/hdr/MarchingCube.h:12)
  approx_extract   <- approximate_isosurfaces_h (isosurfaces_cuda/funcs.h:23-24)
  df_points_*      <- generateHierarchy_h/_W    (isosurfaces_cuda/funcs.h:43-47)
  hist_nmi         <- calculate_histogram_h     (isosurfaces_cuda/funcs.h:32-36)
  field_interval   <- DistanceField::getInterval (types/src/DistanceField.cpp:159-200)
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libiso_amd.so")

OK = 0
ERR_INVALID_ARG = 1
ERR_HIP = 2
ERR_OUT_OF_MEMORY = 3
ERR_CAPACITY = 4
ERR_UNSUPPORTED = 5
ERR_EMPTY = 6

COMPAT_MC = 1 << 0
COMPAT_HIST_CHUNKS = 1 << 1
COMPAT_NMI32 = 1 << 2  # NMI in the reference's float order (else double from exact counts)
COMPAT_ALL = COMPAT_MC | COMPAT_HIST_CHUNKS | COMPAT_NMI32
MC_SORTED = 1 << 8
BATCH_GRAPH = 1 << 9  # iso_ism_batch: pair stage as a captured, replayed HIP graph

DF_LBVH = 0
DF_BRUTE = 1

# every symbol include/iso_amd.h declares (checked by the CPU test suite)
EXPORTS = [
    "iso_ctx_create", "iso_ctx_destroy", "iso_ctx_synchronize", "iso_ctx_stream",
    "iso_ctx_last_error", "iso_ctx_tuning", "iso_status_string", "iso_device_alloc", "iso_device_free",
    "iso_volume_ingest", "iso_equalized_isovalues",
    "iso_copy_to_device", "iso_copy_to_host", "iso_device_memset", "iso_device_copy",
    "iso_mc_extract", "iso_approx_extract", "iso_df_points_f3", "iso_df_points_i3",
    "iso_df_points_normalized", "iso_df_voxel_units", "iso_find_extrema",
    "iso_field_interval", "iso_hist_nmi", "iso_hist_nmi_noalloc", "iso_ism_volume", "iso_ism_batch",
    "iso_gen_blobs", "iso_gen_front", "iso_gen_sines", "iso_gen_timeseries",
    "iso_select_representatives", "iso_select_representatives_ordered", "iso_server_create", "iso_server_add_volume",
    "iso_server_start", "iso_server_submit", "iso_server_wait", "iso_server_get_stats",
    "iso_server_destroy",
]


class IsoError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"[status {status}] {msg}")
        self.status = status


class McStats(C.Structure):
    _fields_ = [("active_cells", C.c_int64), ("total_verts", C.c_int64),
                ("triangles", C.c_int64), ("unique_verts", C.c_int64)]


class IsmStats(C.Structure):
    _fields_ = [("n_nonempty", C.c_int), ("total_points", C.c_int64), ("ms_extract", C.c_double),
                ("ms_df", C.c_double), ("ms_hist", C.c_double), ("graph_replays", C.c_int64),
                ("ms_walk", C.c_double)]


class IsmJob(C.Structure):
    _fields_ = [("d_vol", C.c_void_p), ("X", C.c_int), ("Y", C.c_int), ("Z", C.c_int),
                ("h_isos", C.c_void_p), ("k", C.c_int), ("d_map", C.c_void_p),
                ("h_ids", C.c_void_p), ("n_nonempty", C.c_int)]


EXTRACT_MC = 0
EXTRACT_APPROX = 1

REQ_ISM, REQ_TSM, REQ_VSM = 0, 1, 2


class ServerConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("slots", C.c_int), ("mem_budget", C.c_size_t), ("k", C.c_int),
                ("ds", C.c_int), ("H_ism", C.c_int), ("H_multi", C.c_int), ("reps", C.c_int),
                ("extractor", C.c_int), ("df_method", C.c_int), ("flags", C.c_uint32)]


class Request(C.Structure):
    _fields_ = [("kind", C.c_int), ("var1", C.c_int), ("step1", C.c_int), ("var2", C.c_int),
                ("step2", C.c_int)]


class RequestInfo(C.Structure):
    _fields_ = [("status", C.c_int), ("m", C.c_int), ("slot", C.c_int), ("t_submit", C.c_double),
                ("t_start", C.c_double), ("t_done", C.c_double), ("computed", C.c_int),
                ("computed_pairs", C.c_int64)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("ism_computed", C.c_int64),
                ("cache_hits", C.c_int64), ("evictions", C.c_int64), ("cache_bytes", C.c_size_t),
                ("resident_volumes", C.c_int), ("tunings_in_requests", C.c_int64),
                ("untuned_shapes", C.c_int64)]


_lib = None


def lib() -> C.CDLL:
    """Loads libiso_amd.so once; raises if it is missing (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    vp, i32, i64, u32, u64, f32 = C.c_void_p, C.c_int, C.c_int64, C.c_uint32, C.c_uint64, C.c_float
    sz = C.c_size_t
    sig = {
        "iso_ctx_create": (i32, [C.POINTER(vp), i32]),
        "iso_ctx_destroy": (i32, [vp]),
        "iso_ctx_synchronize": (i32, [vp]),
        "iso_ctx_stream": (vp, [vp]),
        "iso_ctx_last_error": (C.c_char_p, [vp]),
        "iso_ctx_tuning": (C.c_char_p, [vp]),
        "iso_status_string": (C.c_char_p, [i32]),
        "iso_device_alloc": (i32, [vp, C.POINTER(vp), sz]),
        "iso_device_free": (i32, [vp, vp]),
        "iso_copy_to_device": (i32, [vp, vp, vp, sz]),
        "iso_copy_to_host": (i32, [vp, vp, vp, sz]),
        "iso_device_memset": (i32, [vp, vp, i32, sz]),
        "iso_device_copy": (i32, [vp, vp, vp, sz]),
        "iso_mc_extract": (i32, [vp, vp, i32, i32, i32, f32, u32, vp, i64,
                                 C.POINTER(i64), C.POINTER(McStats)]),
        "iso_approx_extract": (i32, [vp, vp, i32, i32, i32, f32, vp, i64, C.POINTER(i64)]),
        "iso_df_points_f3": (i32, [vp, vp, i64, i32, i32, i32, i32, i32, vp]),
        "iso_df_points_i3": (i32, [vp, vp, i64, i32, i32, i32, i32, i32, vp]),
        "iso_df_points_normalized": (i32, [vp, vp, i64, i32, i32, i32, i32, i32, vp]),
        "iso_df_voxel_units": (i32, [vp, vp, i64, i32, i32, i32, i32, vp]),
        "iso_find_extrema": (i32, [vp, vp, i32, i32, i32, vp, i64, C.POINTER(i64)]),
        "iso_field_interval": (i32, [vp, vp, i32, i64, vp, vp]),
        "iso_hist_nmi": (i32, [vp, vp, i32, i64, i32, i32, u32, vp, vp, vp, vp, vp, vp]),
        "iso_hist_nmi_noalloc": (i32, [vp, vp, i32, i64, i32, i32, u32, vp, vp, vp, vp, vp, vp]),
        "iso_ism_volume": (i32, [vp, vp, i32, i32, i32, vp, i32, i32, i32, i32, i32, u32, vp, vp,
                                 vp, C.POINTER(IsmStats)]),
        "iso_ism_batch": (i32, [vp, vp, i32, i32, i32, i32, i32, u32, C.POINTER(IsmStats)]),
        "iso_volume_ingest": (i32, [vp, vp, i32, i32, i32, f32, f32, vp, vp, vp]),
        "iso_equalized_isovalues": (i32, [vp, vp, i64, f32, f32, i32, vp]),
        "iso_select_representatives": (i32, [vp, i32, vp, i32, vp, vp, C.POINTER(i32)]),
        "iso_select_representatives_ordered": (i32, [vp, i32, vp, i32, vp, vp, vp,
                                                     C.POINTER(i32)]),
        "iso_server_create": (i32, [C.POINTER(vp), C.POINTER(ServerConfig)]),
        "iso_server_add_volume": (i32, [vp, i32, i32, vp, i32, i32, i32]),
        "iso_server_start": (i32, [vp]),
        "iso_server_submit": (i32, [vp, C.POINTER(Request), C.POINTER(u64)]),
        "iso_server_wait": (i32, [vp, u64, vp, i32, vp, i32, C.POINTER(RequestInfo)]),
        "iso_server_get_stats": (i32, [vp, C.POINTER(ServerStats)]),
        "iso_server_destroy": (i32, [vp]),
        "iso_gen_timeseries": (i32, [vp, i32, i32, i32, i32, i32, i32, u64]),
        "iso_gen_blobs": (i32, [vp, i32, i32, i32, i32, f32, u64]),
        "iso_gen_front": (i32, [vp, i32, i32, i32, u64]),
        "iso_gen_sines": (i32, [vp, i32, i32, i32, u64]),
    }
    for name, (res, args) in sig.items():
        if hasattr(L, name):
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args


def _np_ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


# ---- host generators (no GPU needed) ---------------------------------------

def gen_blobs(X: int, Y: int, Z: int, nblobs: int = 8, noise: float = 0.01, seed: int = 1) -> np.ndarray:
    img = np.empty((X, Y, Z), np.float32)
    st = lib().iso_gen_blobs(_np_ptr(img), X, Y, Z, nblobs, noise, seed)
    if st != OK:
        raise IsoError(st, "iso_gen_blobs")
    return img


def gen_front(X: int, Y: int, Z: int, seed: int = 2) -> np.ndarray:
    img = np.empty((X, Y, Z), np.float32)
    st = lib().iso_gen_front(_np_ptr(img), X, Y, Z, seed)
    if st != OK:
        raise IsoError(st, "iso_gen_front")
    return img


def gen_sines(X: int, Y: int, Z: int, seed: int = 4) -> np.ndarray:
    img = np.empty((X, Y, Z), np.float32)
    st = lib().iso_gen_sines(_np_ptr(img), X, Y, Z, seed)
    if st != OK:
        raise IsoError(st, "iso_gen_sines")
    return img


def gen_timeseries(X: int, Y: int, Z: int, var: int, step: int, nblobs: int = 10,
                   seed: int = 5) -> np.ndarray:
    img = np.empty((X, Y, Z), np.float32)
    st = lib().iso_gen_timeseries(_np_ptr(img), X, Y, Z, var, step, nblobs, seed)
    if st != OK:
        raise IsoError(st, "iso_gen_timeseries")
    return img


def select_representatives(smap: np.ndarray, cnt: int, ids=None):
    """Representative ids (ascending) and their priorities (host)."""
    smap = np.ascontiguousarray(smap, np.float32)
    m = smap.shape[0]
    ids_arr = None if ids is None else np.ascontiguousarray(ids, np.int32)
    out = np.zeros(min(cnt, m), np.int32)
    pri = np.zeros(min(cnt, m), np.float32)
    got = C.c_int(0)
    st = lib().iso_select_representatives(_np_ptr(smap), m,
                                          None if ids_arr is None else _np_ptr(ids_arr), cnt,
                                          _np_ptr(out), _np_ptr(pri), C.byref(got))
    if st != OK:
        raise IsoError(st, "iso_select_representatives")
    return out[: got.value], pri[: got.value]


def select_representatives_ordered(smap: np.ndarray, cnt: int, ids=None):
    """Representatives in the order they are taken: (ids, map rows, priorities)."""
    smap = np.ascontiguousarray(smap, np.float32)
    m = smap.shape[0]
    ids_arr = None if ids is None else np.ascontiguousarray(ids, np.int32)
    n = min(cnt, m)
    sel, rows, pri = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(n, np.float32)
    got = C.c_int(0)
    st = lib().iso_select_representatives_ordered(
        _np_ptr(smap), m, None if ids_arr is None else _np_ptr(ids_arr), cnt, _np_ptr(sel),
        _np_ptr(rows), _np_ptr(pri), C.byref(got))
    if st != OK:
        raise IsoError(st, "iso_select_representatives_ordered")
    g = got.value
    return sel[:g], rows[:g], pri[:g]


class Server:
    """Serving form (iso_server_*): resident volumes, ISM/TSM/VSM requests."""

    def __init__(self, *, device=0, slots=4, mem_budget=64 << 30, k=64, ds=2, H_ism=128,
                 H_multi=128, reps=16, extractor=EXTRACT_MC, df_method=DF_LBVH, flags=COMPAT_ALL):
        lib()
        self.cfg = ServerConfig(device, slots, mem_budget, k, ds, H_ism, H_multi, reps, extractor,
                                df_method, flags)
        self.h = C.c_void_p()
        st = lib().iso_server_create(C.byref(self.h), C.byref(self.cfg))
        if st != OK:
            raise IsoError(st, "iso_server_create")

    def add_volume(self, var: int, step: int, img: np.ndarray):
        img = np.ascontiguousarray(img, np.float32)
        st = lib().iso_server_add_volume(self.h, var, step, _np_ptr(img), *img.shape)
        if st != OK:
            raise IsoError(st, "iso_server_add_volume")

    def start(self):
        st = lib().iso_server_start(self.h)
        if st != OK:
            raise IsoError(st, "iso_server_start")

    def submit(self, kind: int, var1: int, step1: int, var2: int = 0, step2: int = 0) -> int:
        t = C.c_uint64()
        st = lib().iso_server_submit(self.h, C.byref(Request(kind, var1, step1, var2, step2)),
                                     C.byref(t))
        if st != OK:
            raise IsoError(st, "iso_server_submit")
        return t.value

    def wait(self, ticket: int):
        """(map m x m, ids, RequestInfo)."""
        k = self.cfg.k
        cap = max(k * k, 4 * self.cfg.reps * self.cfg.reps)
        buf = np.zeros(cap, np.float32)
        ids = np.zeros(max(k, 2 * self.cfg.reps), np.int32)
        info = RequestInfo()
        st = lib().iso_server_wait(self.h, ticket, _np_ptr(buf), cap, _np_ptr(ids), ids.size,
                                   C.byref(info))
        if st != OK:
            raise IsoError(st, f"iso_server_wait({ticket})")
        m = info.m
        return buf[: m * m].reshape(m, m).copy(), ids.copy(), info

    def stats(self) -> ServerStats:
        s = ServerStats()
        lib().iso_server_get_stats(self.h, C.byref(s))
        return s

    def close(self):
        if self.h:
            lib().iso_server_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class DeviceArray:
    """Device buffer owned by a Context (freed with it or by .free())."""

    def __init__(self, ctx: "Context", ptr: int, shape, dtype):
        self.ctx, self.ptr, self.shape, self.dtype = ctx, ptr, tuple(shape), np.dtype(dtype)

    @property
    def nbytes(self) -> int:
        return int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize

    def offset(self, elements: int) -> int:
        return self.ptr + elements * self.dtype.itemsize

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        self.ctx._check(lib().iso_copy_to_host(self.ctx.h, _np_ptr(out), C.c_void_p(self.ptr), self.nbytes))
        return out

    def free(self):
        if self.ptr:
            self.ctx._check(lib().iso_device_free(self.ctx.h, C.c_void_p(self.ptr)))
            self.ctx._owned.discard(self.ptr)
            self.ptr = 0


class Context:
    """One HIP device + stream + workspace (iso_ctx)."""

    def __init__(self, device: int = 0):
        self.h = C.c_void_p()
        st = lib().iso_ctx_create(C.byref(self.h), device)
        if st != OK:
            raise IsoError(st, f"iso_ctx_create(device={device})")
        self._owned = set()

    def close(self):
        if self.h:
            for p in list(self._owned):
                lib().iso_device_free(self.h, C.c_void_p(p))
            self._owned.clear()
            lib().iso_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, st: int, what: str = ""):
        if st != OK:
            msg = lib().iso_ctx_last_error(self.h).decode()
            raise IsoError(st, f"{what}: {lib().iso_status_string(st).decode()} {msg}")

    def synchronize(self):
        self._check(lib().iso_ctx_synchronize(self.h), "synchronize")

    def tuning(self) -> list:
        """The autotuner's choices so far (one line per shape class)."""
        return [l for l in lib().iso_ctx_tuning(self.h).decode().splitlines() if l]

    @property
    def stream(self) -> int:
        return lib().iso_ctx_stream(self.h) or 0

    # ---- memory ----
    def empty(self, shape, dtype=np.float32) -> DeviceArray:
        shape = tuple(int(s) for s in np.atleast_1d(shape))
        n = int(np.prod(shape, dtype=np.int64)) * np.dtype(dtype).itemsize
        p = C.c_void_p()
        self._check(lib().iso_device_alloc(self.h, C.byref(p), max(n, 4)), "alloc")
        self._owned.add(p.value)
        return DeviceArray(self, p.value, shape, dtype)

    def zeros(self, shape, dtype=np.float32) -> DeviceArray:
        d = self.empty(shape, dtype)
        self._check(lib().iso_device_memset(self.h, C.c_void_p(d.ptr), 0, d.nbytes), "memset")
        return d

    def to_device(self, a: np.ndarray) -> DeviceArray:
        a = np.ascontiguousarray(a)
        d = self.empty(a.shape, a.dtype)
        self._check(lib().iso_copy_to_device(self.h, C.c_void_p(d.ptr), _np_ptr(a), a.nbytes), "h2d")
        return d

    def copy(self, dst: DeviceArray, src: DeviceArray, nbytes: int | None = None):
        """Device-to-device copy of nbytes (default: all of src)."""
        n = src.nbytes if nbytes is None else int(nbytes)
        if n > dst.nbytes or n > src.nbytes:
            raise ValueError("copy larger than a buffer")
        self._check(lib().iso_device_copy(self.h, C.c_void_p(dst.ptr), C.c_void_p(src.ptr), n),
                    "d2d")

    # ---- ingest / isovalues ----
    def volume_ingest(self, d_file: DeviceArray, dims, cfg_min: float, cfg_max: float):
        """x-fastest buffer -> z-fastest image; returns (image, real_min, real_max)."""
        X, Y, Z = (int(d) for d in dims)
        img = self.empty((X, Y, Z), np.float32)
        lo, hi = C.c_float(), C.c_float()
        self._check(lib().iso_volume_ingest(self.h, C.c_void_p(d_file.ptr), X, Y, Z,
                                            C.c_float(cfg_min), C.c_float(cfg_max),
                                            C.c_void_p(img.ptr), C.byref(lo), C.byref(hi)),
                    "iso_volume_ingest")
        return img, lo.value, hi.value

    def equalized_isovalues(self, img: DeviceArray, real_min: float, real_max: float, k: int):
        out = np.zeros(k, np.float32)
        n = int(np.prod(img.shape))
        self._check(lib().iso_equalized_isovalues(self.h, C.c_void_p(img.ptr), n,
                                                  C.c_float(real_min), C.c_float(real_max), k,
                                                  _np_ptr(out)), "iso_equalized_isovalues")
        return out

    # ---- hot path ----
    def mc_extract(self, vol: DeviceArray, dims, iso: float, flags: int = COMPAT_MC,
                   out: DeviceArray | None = None):
        """Unique marching-cubes vertices; returns (DeviceArray n x 3, count, McStats)."""
        gx, gy, gz = (int(d) for d in dims)
        cnt = C.c_int64(0)
        stats = McStats()
        cap = out.shape[0] if out is not None else 0
        optr = out.ptr if out is not None else None
        st = lib().iso_mc_extract(self.h, C.c_void_p(vol.ptr), gx, gy, gz, C.c_float(iso), flags,
                                  C.c_void_p(optr), cap, C.byref(cnt), C.byref(stats))
        if st == ERR_CAPACITY:
            out = self.empty((max(cnt.value, 1), 3), np.float32)
            st = lib().iso_mc_extract(self.h, C.c_void_p(vol.ptr), gx, gy, gz, C.c_float(iso), flags,
                                      C.c_void_p(out.ptr), out.shape[0], C.byref(cnt), C.byref(stats))
        self._check(st, "iso_mc_extract")
        return out, cnt.value, stats

    def approx_extract(self, img: DeviceArray, dims, iso: float, out: DeviceArray | None = None):
        X, Y, Z = (int(d) for d in dims)
        cnt = C.c_int64(0)
        cap = out.shape[0] if out is not None else 0
        optr = out.ptr if out is not None else None
        st = lib().iso_approx_extract(self.h, C.c_void_p(img.ptr), X, Y, Z, C.c_float(iso),
                                      C.c_void_p(optr), cap, C.byref(cnt))
        if st == ERR_CAPACITY:
            out = self.empty((max(cnt.value, 1), 3), np.int32)
            st = lib().iso_approx_extract(self.h, C.c_void_p(img.ptr), X, Y, Z, C.c_float(iso),
                                          C.c_void_p(out.ptr), out.shape[0], C.byref(cnt))
        self._check(st, "iso_approx_extract")
        return out, cnt.value

    def df_points(self, pts: DeviceArray, n: int, dims, ds: int, method: int = DF_LBVH,
                  out: DeviceArray | None = None) -> DeviceArray:
        X, Y, Z = (int(d) for d in dims)
        F = (X // ds) * (Y // ds) * (Z // ds)
        if out is None:
            out = self.empty((F,), np.float32)
        fn = lib().iso_df_points_i3 if pts.dtype == np.int32 else lib().iso_df_points_f3
        self._check(fn(self.h, C.c_void_p(pts.ptr), n, X, Y, Z, ds, method, C.c_void_p(out.ptr)),
                    "iso_df_points")
        return out

    def df_voxel_units(self, pts: DeviceArray, n: int, grid, ds: int,
                       out: DeviceArray | None = None) -> DeviceArray:
        """calculate_distances_h: brute force in voxel units over an
        nx*ny*nz grid of samples at (x*ds + ds/2, ...)."""
        nx, ny, nz = (int(d) for d in grid)
        if out is None:
            out = self.empty((nx * ny * nz,), np.float32)
        self._check(lib().iso_df_voxel_units(self.h, C.c_void_p(pts.ptr if pts else None), n, nx,
                                             ny, nz, ds, C.c_void_p(out.ptr)),
                    "iso_df_voxel_units")
        return out

    def find_extrema(self, vol: DeviceArray, dims) -> np.ndarray:
        """findExtrema_h: the values of the volume's local extrema, voxel order."""
        X, Y, Z = (int(d) for d in dims)
        cnt = C.c_int64(0)
        st = lib().iso_find_extrema(self.h, C.c_void_p(vol.ptr), X, Y, Z, None, 0, C.byref(cnt))
        if st == ERR_CAPACITY:
            out = self.empty((max(cnt.value, 1),), np.float32)
            st = lib().iso_find_extrema(self.h, C.c_void_p(vol.ptr), X, Y, Z, C.c_void_p(out.ptr),
                                        out.shape[0], C.byref(cnt))
            self._check(st, "iso_find_extrema")
            vals = out.numpy()[: cnt.value].copy()
            out.free()
            return vals
        self._check(st, "iso_find_extrema")
        return np.zeros(0, np.float32)

    def field_interval(self, fields: DeviceArray, nfields: int, F: int):
        mins = self.empty((nfields,), np.float32)
        maxs = self.empty((nfields,), np.float32)
        self._check(lib().iso_field_interval(self.h, C.c_void_p(fields.ptr), nfields, F,
                                             C.c_void_p(mins.ptr), C.c_void_p(maxs.ptr)),
                    "iso_field_interval")
        return mins, maxs

    def hist_nmi(self, fields: DeviceArray, n: int, F: int, H: int, multi: bool = False,
                 flags: int = COMPAT_ALL, mins: DeviceArray | None = None,
                 maxs: DeviceArray | None = None, want_joint: bool = False,
                 no_alloc: bool = False):
        """Similarity map (n x n float32); optionally joint histograms + marginals.
        no_alloc: through iso_hist_nmi_noalloc (mins and maxs must be given,
        the outputs are allocated here, not by the call)."""
        if mins is None or maxs is None:
            mins, maxs = self.field_interval(fields, 2 * n if multi else n, F)
        smap = self.empty((n, n), np.float32)
        joint = rows = cols = None
        if want_joint:
            joint = self.empty((n, n, H, H), np.uint32)
            rows = self.empty((n, n, H), np.uint32)
            cols = self.empty((n, n, H), np.uint32)
        fn = lib().iso_hist_nmi_noalloc if no_alloc else lib().iso_hist_nmi
        self._check(fn(
            self.h, C.c_void_p(fields.ptr), n, F, H, int(multi), flags, C.c_void_p(mins.ptr),
            C.c_void_p(maxs.ptr), C.c_void_p(smap.ptr),
            C.c_void_p(joint.ptr if joint else None), C.c_void_p(rows.ptr if rows else None),
            C.c_void_p(cols.ptr if cols else None)),
            "iso_hist_nmi_noalloc" if no_alloc else "iso_hist_nmi")
        return smap, joint, rows, cols

    def ism_volume(self, vol: DeviceArray, dims, isos, extractor: int = EXTRACT_MC, ds: int = 1,
                   H: int = 128, df_method: int = DF_LBVH, flags: int = COMPAT_ALL,
                   fields: DeviceArray | None = None, smap: DeviceArray | None = None,
                   want_fields: bool = False):
        """Batched ISM of one volume.  Returns (map DeviceArray (k*k, first m*m
        used), surface ids (m,), IsmStats, fields DeviceArray or None)."""
        X, Y, Z = (int(d) for d in dims)
        isos = np.ascontiguousarray(isos, np.float32)
        k = int(isos.size)
        F = (X // ds) * (Y // ds) * (Z // ds)
        if smap is None:
            smap = self.empty((k * k,), np.float32)
        if fields is None and want_fields:
            fields = self.empty((k, F), np.float32)
        ids = np.zeros(k, np.int32)
        st = IsmStats()
        self._check(lib().iso_ism_volume(
            self.h, C.c_void_p(vol.ptr), X, Y, Z, _np_ptr(isos), k, extractor, ds, H, df_method,
            flags, C.c_void_p(smap.ptr), C.c_void_p(fields.ptr if fields else None), _np_ptr(ids),
            C.byref(st)), "iso_ism_volume")
        return smap, ids[: st.n_nonempty].copy(), st, fields

    def ism_batch(self, vols, dims_list, isos_list, extractor: int = EXTRACT_MC, ds: int = 1,
                  H: int = 128, df_method: int = DF_LBVH, flags: int = COMPAT_ALL):
        """ISMs of several volumes in one call (one persistent pair launch).
        Returns ([(map m x m numpy, ids (m,))] per volume, IsmStats)."""
        n = len(vols)
        jobs = (IsmJob * n)()
        keep = []
        for b in range(n):
            X, Y, Z = (int(d) for d in dims_list[b])
            isos = np.ascontiguousarray(isos_list[b], np.float32)
            k = int(isos.size)
            smap = self.empty((k * k,), np.float32)
            ids = np.zeros(k, np.int32)
            keep.append((isos, smap, ids))
            jobs[b] = IsmJob(vols[b].ptr, X, Y, Z, isos.ctypes.data, k, smap.ptr, ids.ctypes.data, 0)
        st = IsmStats()
        self._check(lib().iso_ism_batch(self.h, C.cast(jobs, C.c_void_p), n, extractor, ds, H,
                                        df_method, flags, C.byref(st)), "iso_ism_batch")
        out = []
        for b in range(n):
            m = jobs[b].n_nonempty
            isos, smap, ids = keep[b]
            out.append((smap.numpy()[: m * m].reshape(m, m), ids[:m].copy()))
        return out, st
