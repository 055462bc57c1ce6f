"""iso_hist_nmi_noalloc (no device allocation; working memory in a fixed
per-device array) against iso_hist_nmi on the same inputs: the same maps,
joint counts and marginals, bit for bit, on two calls in a row -- the second
smaller -- in both NMI modes, and sizes past its maximum rejected."""
import numpy as np
import pytest

import isoamd

pytestmark = pytest.mark.gpu


def _fields(n, F, seed):
    rng = np.random.default_rng(seed)
    base = rng.random(F).astype(np.float32)
    return np.stack([np.abs(base - 0.04 * k + 0.03 * rng.standard_normal(F)).astype(np.float32)
                     for k in range(n)])


@pytest.mark.parametrize("flags", [isoamd.COMPAT_ALL, isoamd.COMPAT_HIST_CHUNKS],
                         ids=["nmi32", "nmi64"])
def test_no_alloc_equals_allocating_entry_point_twice(flags):
    # first call: 300 pairs of 256-bin histograms, more than one launch's
    # worth of the fixed array; second call: fewer, smaller fields, 128 bins
    with isoamd.Context(0) as ctx:
        for n, F, H, multi, seed in ((24, 60000, 256, False, 1), (5, 20000, 128, True, 2)):
            print(f"no-alloc: n={n} F={F} H={H} multi={multi} flags={flags}", flush=True)
            nf = 2 * n if multi else n
            f = ctx.to_device(_fields(nf, F, seed))
            mins, maxs = ctx.field_interval(f, nf, F)
            want = ctx.hist_nmi(f, n, F, H, multi=multi, flags=flags, mins=mins, maxs=maxs,
                                want_joint=True)
            got = ctx.hist_nmi(f, n, F, H, multi=multi, flags=flags, mins=mins, maxs=maxs,
                               want_joint=True, no_alloc=True)
            for g, w in zip(got, want):
                assert np.array_equal(g.numpy(), w.numpy())


def test_no_alloc_rejects_sizes_past_its_maximum():
    import ctypes as C

    with isoamd.Context(0) as ctx:
        f = ctx.to_device(_fields(2, 1000, 3))
        mins, maxs = ctx.field_interval(f, 2, 1000)
        smap = ctx.empty((4,), np.float32)
        for n, H in ((2, 257), (70000, 16)):  # rejected before any output is touched
            st = isoamd.lib().iso_hist_nmi_noalloc(
                ctx.h, C.c_void_p(f.ptr), n, 1000, H, 0, isoamd.COMPAT_ALL, C.c_void_p(mins.ptr),
                C.c_void_p(maxs.ptr), C.c_void_p(smap.ptr), None, None, None)
            assert st == isoamd.ERR_UNSUPPORTED
            assert b"noalloc" in isoamd.lib().iso_ctx_last_error(ctx.h)
