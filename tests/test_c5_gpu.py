"""The serving workload's shape: 256x256x128 volumes of the multivariate
time-varying set (3 variables; three time steps here), ISMs of 64 isovalues,
16 representatives per volume, ds=2, 128 bins.

Each request gives the same result alone (one slot, one request at a time)
and in a concurrent mix on four slots.  The ISM of a volume equals the
single-volume pipeline's; a VSM and a TSM equal the oracle's reference-order
cross map (the `multi` task list, JointHistogram.cu:309-318) over the
representative fields; those fields equal the oracle's on sampled voxels."""
import numpy as np
import pytest

import isoamd
import oracle

pytestmark = pytest.mark.gpu

DIMS, K, DS, H, R = (256, 256, 128), 64, 2, 128, 16
F = (DIMS[0] // DS) * (DIMS[1] // DS) * (DIMS[2] // DS)
REQS = [(isoamd.REQ_ISM, 0, 0, 0, 0), (isoamd.REQ_TSM, 0, 0, 0, 1), (isoamd.REQ_VSM, 0, 1, 1, 1),
        (isoamd.REQ_ISM, 2, 2, 0, 0), (isoamd.REQ_TSM, 1, 1, 1, 2), (isoamd.REQ_VSM, 1, 2, 2, 2),
        (isoamd.REQ_ISM, 1, 0, 0, 0), (isoamd.REQ_VSM, 0, 2, 2, 2)]


@pytest.fixture(scope="module")
def vols():
    return {(v, t): isoamd.gen_timeseries(*DIMS, v, t, nblobs=10, seed=5)
            for v in range(3) for t in range(3)}


def _server(vols, slots):
    s = isoamd.Server(slots=slots, k=K, ds=DS, H_ism=H, H_multi=H, reps=R, mem_budget=32 << 30)
    for (v, t), img in vols.items():
        s.add_volume(v, t, img)
    s.start()
    return s


def _isos(img):
    lo, hi = float(img.min()), float(img.max())
    return np.array([lo + (i + 1) * (np.float64(hi) - lo) / (K + 1) for i in range(K)], np.float32)


def test_c5_alone_equals_mixed(vols):
    print("c5: requests one at a time", flush=True)
    with _server(vols, 1) as s:
        alone = [s.wait(s.submit(*r))[:2] for r in REQS]
    print("c5: the same requests twice over, concurrently on four slots", flush=True)
    with _server(vols, 4) as s:
        tickets = [s.submit(*r) for r in REQS + REQS[::-1]]
        mixed = [s.wait(t) for t in tickets]
        assert len({info.slot for _, _, info in mixed}) > 1
    for q, (mp, ids, info) in enumerate(mixed):
        r = q if q < len(REQS) else 2 * len(REQS) - 1 - q
        assert info.status == 0
        assert np.array_equal(mp, alone[r][0]) and np.array_equal(ids, alone[r][1])


def test_c5_maps_against_oracle(ctx, vols):
    def reps(key):
        img = vols[key]
        smap, ids, st, fields = ctx.ism_volume(ctx.to_device(img), DIMS, _isos(img),
                                               isoamd.EXTRACT_MC, DS, H, want_fields=True)
        m = st.n_nonempty
        mp = smap.numpy()[: m * m].reshape(m, m)
        rep, _ = isoamd.select_representatives(mp, R)
        return mp, ids, fields.numpy()[:m], rep

    print("c5: single-volume pipelines of (0,1), (1,1), (0,2)", flush=True)
    mA, idsA, fA, rA = reps((0, 1))
    _, idsB, fB, rB = reps((1, 1))
    _, idsC, fC, rC = reps((0, 2))
    with _server(vols, 2) as s:
        ism, iids, _ = s.wait(s.submit(isoamd.REQ_ISM, 0, 1))
        vsm, vids, _ = s.wait(s.submit(isoamd.REQ_VSM, 0, 1, 1, 1))
        tsm, tids, _ = s.wait(s.submit(isoamd.REQ_TSM, 0, 1, 0, 2))
    assert np.array_equal(ism, mA) and list(iids[: len(idsA)]) == list(idsA)
    print("c5: oracle cross maps over the representative fields", flush=True)
    for got, ids, fb, idsb, rb in ((vsm, vids, fB, idsB, rB), (tsm, tids, fC, idsC, rC)):
        assert list(ids[: 2 * R]) == list(idsA[rA]) + list(idsb[rb])
        m32, _ = oracle.ism(np.concatenate([fA[rA], fb[rb]]), H, multi=True)
        assert np.array_equal(got, m32)
    # the representative fields themselves against the oracle on sampled voxels
    img = vols[(0, 1)]
    r0 = int(rA[0])
    pts, _ = oracle.mc(img.ravel(), DIMS, _isos(img)[idsA[r0]],
                       ncls=oracle.mc_ncls_compat(*DIMS))
    idx = np.sort(np.random.default_rng(1).choice(F, 2048, replace=False)).astype(np.int64)
    assert np.array_equal(fA[r0][idx], oracle.df_tree(oracle.normalize(pts, DIMS), DIMS, DS,
                                                       idx=idx))
