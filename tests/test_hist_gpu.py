"""HIP joint histograms + NMI against the oracle: counts bit-exact; NMI in the
reference's float order (ISO_COMPAT_NMI32, part of COMPAT_ALL) bit-exact
against the oracle's restatement of calculateMutualInformation
(JointHistogram.cu:24-60); NMI in double (without the flag) within 1e-6
relative of the oracle's double evaluation."""
import numpy as np
import pytest

import isoamd
import oracle

pytestmark = pytest.mark.gpu

NMI_RTOL = 1e-6
NMI_ATOL = 1e-7
DOUBLE = isoamd.COMPAT_HIST_CHUNKS  # compat chunks, NMI in double


def _df_like(n, F, seed):
    """Correlated non-negative fields, like distance fields of nearby isovalues."""
    rng = np.random.default_rng(seed)
    base = rng.random(F).astype(np.float32)
    out = []
    for k in range(n):
        f = np.abs(base - 0.05 * k + 0.02 * rng.standard_normal(F).astype(np.float32))
        out.append(f.astype(np.float32))
    return np.stack(out)


def _oracle_hists(fields, H, multi=False, compat=True):
    """(joint counts, double map, reference-order float map) of every ordered pair."""
    nf, F = fields.shape
    n = nf // 2 if multi else nf
    iv = [oracle.interval(f) for f in fields]
    processed = oracle.hist_processed(F) if compat else F
    joint = np.zeros((n, n, H, H), np.uint32)
    m64 = np.zeros((n, n))
    m32 = np.zeros((n, n), np.float32)
    for i in range(n):
        for j in range(n):
            b = n + j if multi else j
            h = oracle.hist_pair(fields[i], fields[b], H, *iv[i], *iv[b], processed=processed)
            joint[i, j] = h
            m64[i, j] = oracle.nmi_f64(h, F)
            m32[i, j] = oracle.nmi_ref32(h, F)
    return joint, m64, m32


def _check(smap, m64, m32, flags):
    got = smap.numpy()
    if flags & isoamd.COMPAT_NMI32:
        assert np.array_equal(got, m32), (
            f"float-order NMI differs from the reference order in "
            f"{int(np.count_nonzero(got != m32))} entries")
    else:
        np.testing.assert_allclose(got, m64, rtol=NMI_RTOL, atol=NMI_ATOL)


@pytest.mark.parametrize("flags", [isoamd.COMPAT_ALL, DOUBLE], ids=["nmi32", "nmi64"])
@pytest.mark.parametrize("n,F,H", [(6, 40000, 128), (5, 30011, 256), (4, 20000, 3),
                                   (3, 50000, 1), (1, 10000, 64), (12, 262144, 128)])
def test_hist_nmi_matches_oracle(ctx, n, F, H, flags):
    print(f"hist parity n={n} F={F} H={H} flags={flags}", flush=True)
    fields = _df_like(n, F, seed=n * 1000 + H)
    d = ctx.to_device(fields)
    smap, joint, rows, cols = ctx.hist_nmi(d, n, F, H, flags=flags, want_joint=True)
    j_ref, m64, m32 = _oracle_hists(fields, H)
    j_gpu = joint.numpy()
    assert np.array_equal(j_gpu, j_ref)
    assert np.array_equal(rows.numpy(), j_ref.sum(3, dtype=np.uint32))
    assert np.array_equal(cols.numpy(), j_ref.sum(2, dtype=np.uint32))
    _check(smap, m64, m32, flags)


def test_nmi32_transposed_entries_follow_their_own_order(ctx):
    # M[i][j] and M[j][i] are separate float chains in the reference (its hXY
    # sums the (j, i) histogram row by row); fields with many populated bins
    # make the two differ in the last bits for some pairs
    n, F, H = 8, 120000, 256
    rng = np.random.default_rng(12)
    fields = np.abs(rng.standard_normal((n, F))).astype(np.float32)
    d = ctx.to_device(fields)
    smap, _, _, _ = ctx.hist_nmi(d, n, F, H)
    _, _, m32 = _oracle_hists(fields, H)
    got = smap.numpy()
    assert np.array_equal(got, m32)
    assert np.count_nonzero(m32 != m32.T) > 0  # the case the test is about


def test_interval_matches_oracle(ctx):
    fields = _df_like(7, 123457, seed=9)
    fields[3] = 0.0  # max stays FLT_MIN (DistanceField.cpp:162)
    d = ctx.to_device(fields)
    mins, maxs = ctx.field_interval(d, 7, fields.shape[1])
    ref = np.array([oracle.interval(f) for f in fields], np.float32)
    assert np.array_equal(mins.numpy(), ref[:, 0]) and np.array_equal(maxs.numpy(), ref[:, 1])


@pytest.mark.parametrize("flags", [isoamd.COMPAT_ALL, DOUBLE], ids=["nmi32", "nmi64"])
def test_multi_cross_map(ctx, flags):
    n, F, H = 4, 25000, 128
    fields = _df_like(2 * n, F, seed=77)
    d = ctx.to_device(fields)
    smap, joint, _, _ = ctx.hist_nmi(d, n, F, H, multi=True, flags=flags, want_joint=True)
    j_ref, m64, m32 = _oracle_hists(fields, H, multi=True)
    assert np.array_equal(joint.numpy(), j_ref)
    _check(smap, m64, m32, flags)


@pytest.mark.parametrize("flags", [isoamd.COMPAT_ALL, DOUBLE], ids=["nmi32", "nmi64"])
def test_16bit_counters_drain_without_wrapping(ctx, flags):
    # one bin receives ~190k counts: the 16-bit LDS counters must drain to 32 bits
    n, F, H = 2, 200000, 256
    rng = np.random.default_rng(5)
    fields = np.zeros((n, F), np.float32)
    fields[:, : F // 20] = rng.random((n, F // 20)).astype(np.float32)
    d = ctx.to_device(fields)
    smap, joint, _, _ = ctx.hist_nmi(d, n, F, H, flags=flags, want_joint=True)
    j_ref, m64, m32 = _oracle_hists(fields, H)
    assert j_ref.max() > 65535
    assert np.array_equal(joint.numpy(), j_ref)
    _check(smap, m64, m32, flags)


def test_nmi32_many_nonzero_bins_use_several_windows(ctx):
    # nearly every one of the 256 x 256 bins populated: the chain lists pass
    # through LDS in several windows
    n, F, H = 3, 400000, 256
    rng = np.random.default_rng(21)
    fields = rng.random((n, F)).astype(np.float32)
    d = ctx.to_device(fields)
    smap, joint, _, _ = ctx.hist_nmi(d, n, F, H, want_joint=True)
    j_ref, _, m32 = _oracle_hists(fields, H)
    assert (j_ref[0, 1] > 0).sum() > 3 * 8192
    assert np.array_equal(joint.numpy(), j_ref)
    assert np.array_equal(smap.numpy(), m32)


@pytest.mark.parametrize("flags", [isoamd.COMPAT_ALL, DOUBLE], ids=["nmi32", "nmi64"])
def test_constant_and_identical_fields(ctx, flags):
    F = 5000
    fields = np.stack([np.full(F, 0.25, np.float32), np.linspace(0, 1, F, dtype=np.float32),
                       np.linspace(0, 1, F, dtype=np.float32)])
    d = ctx.to_device(fields)
    smap, _, _, _ = ctx.hist_nmi(d, 3, F, 128, flags=flags)
    m = smap.numpy()
    _, m64, m32 = _oracle_hists(fields, 128)
    _check(smap, m64, m32, flags)
    if not flags & isoamd.COMPAT_NMI32:
        assert m[0, 0] == 0.0  # zero entropy -> NaN -> 0
    assert m[1, 1] == 1.0 and m[2, 2] == 1.0 and m[1, 2] == 1.0


def test_chunk_quirk_matches_reference(ctx):
    # F = 1,000,000 under the reference's chunk loop: nothing is binned
    F = 1000000
    fields = _df_like(3, F, seed=3)
    d = ctx.to_device(fields)
    smap, _, _, _ = ctx.hist_nmi(d, 3, F, 32)
    assert np.all(smap.numpy() == 1.0)
    smap2, _, _, _ = ctx.hist_nmi(d, 3, F, 32, flags=0)
    _, m64, _ = _oracle_hists(fields, 32, compat=False)
    np.testing.assert_allclose(smap2.numpy(), m64, rtol=NMI_RTOL, atol=NMI_ATOL)
