"""Serving form: requests give the same maps alone and in a concurrent mix,
cross maps equal the batched pipeline + representatives + multi histogram,
memory admission evicts without changing results; representative selection
properties on the CPU."""
import numpy as np
import pytest

import isoamd

DIMS = (48, 40, 32)
K, DS, H, R = 12, 2, 64, 4


def test_representatives_pick_one_per_cluster():
    # three clusters of mutually similar isosurfaces
    m = 9
    smap = np.full((m, m), 0.1, np.float32)
    for c in range(3):
        idx = slice(3 * c, 3 * c + 3)
        smap[idx, idx] = 0.9
    np.fill_diagonal(smap, 1.0)
    ids, pri = isoamd.select_representatives(smap, 3)
    assert sorted(set(i // 3 for i in ids)) == [0, 1, 2]
    assert list(ids) == sorted(ids)
    ids2, _ = isoamd.select_representatives(smap, 3)
    assert np.array_equal(ids, ids2)


def test_representatives_count_and_mapping():
    rng = np.random.default_rng(0)
    a = rng.random((7, 7)).astype(np.float32)
    smap = (a + a.T) / 2
    np.fill_diagonal(smap, 1.0)
    ids, _ = isoamd.select_representatives(smap, 16, ids=np.arange(7) * 10)
    assert len(ids) == 7 and set(ids) == set(range(0, 70, 10))  # cnt > m takes all


def _volumes():
    vols = {}
    for var in range(2):
        for step in range(3):
            vols[(var, step)] = isoamd.gen_timeseries(*DIMS, var, step, nblobs=6, seed=9)
    return vols


REQS = [(isoamd.REQ_ISM, 0, 0, 0, 0), (isoamd.REQ_TSM, 0, 0, 0, 1), (isoamd.REQ_VSM, 0, 1, 1, 1),
        (isoamd.REQ_ISM, 1, 2, 0, 0), (isoamd.REQ_TSM, 1, 1, 1, 2), (isoamd.REQ_VSM, 0, 2, 1, 2),
        (isoamd.REQ_ISM, 0, 1, 0, 0), (isoamd.REQ_TSM, 0, 1, 0, 2)]


def _server(vols, **kw):
    s = isoamd.Server(k=K, ds=DS, H_ism=H, H_multi=H, reps=R, **kw)
    for (v, t), img in vols.items():
        s.add_volume(v, t, img)
    s.start()
    return s


@pytest.mark.gpu
def test_alone_equals_mixed():
    vols = _volumes()
    with _server(vols, slots=1) as s:
        alone = []
        for r in REQS:
            mp, ids, info = s.wait(s.submit(*r))
            assert info.status == 0
            alone.append((mp, ids))
    with _server(vols, slots=4) as s:
        tickets = [s.submit(*r) for r in REQS + REQS]
        mixed = [s.wait(t) for t in tickets]
        st = s.stats()
        touched = {(r[1], r[2]) for r in REQS} | {(r[3], r[4]) for r in REQS if r[0] != 0}
        assert len(touched) == 5  # volume (1, 0) is never asked for
        # each touched volume's ISM is computed once, however many requests share it
        assert st.completed == 2 * len(REQS) and st.ism_computed == len(touched)
        assert len({info.slot for _, _, info in mixed}) > 1  # work spread over slots
    for i, (mp, ids, info) in enumerate(mixed):
        a_mp, a_ids = alone[i % len(REQS)]
        assert np.array_equal(mp, a_mp) and np.array_equal(ids, a_ids)
        assert info.t_submit <= info.t_start <= info.t_done


@pytest.mark.gpu
def test_cross_map_matches_pipeline(ctx):
    vols = _volumes()
    F = (DIMS[0] // DS) * (DIMS[1] // DS) * (DIMS[2] // DS)

    def reps_fields(img):
        lo, hi = float(img.min()), float(img.max())
        isos = np.array([lo + (i + 1) * (np.float64(hi) - lo) / (K + 1) for i in range(K)], np.float32)
        smap, ids, st, fields = ctx.ism_volume(ctx.to_device(img), DIMS, isos, isoamd.EXTRACT_MC, DS,
                                               H, want_fields=True)
        m = st.n_nonempty
        rep, _ = isoamd.select_representatives(smap.numpy()[: m * m].reshape(m, m), R)
        return smap.numpy()[: m * m].reshape(m, m), ids, fields.numpy()[rep], ids[rep]

    mA, idsA, fA, rA = reps_fields(vols[(0, 1)])
    _, _, fB, rB = reps_fields(vols[(1, 1)])
    both = ctx.to_device(np.concatenate([fA, fB]))
    want, _, _, _ = ctx.hist_nmi(both, R, F, H, multi=True)
    with _server(vols, slots=2) as s:
        got, ids, info = s.wait(s.submit(isoamd.REQ_VSM, 0, 1, 1, 1))
        ism, iids, _ = s.wait(s.submit(isoamd.REQ_ISM, 0, 1))
    assert np.array_equal(got, want.numpy())
    assert list(ids[: 2 * R]) == list(rA) + list(rB)
    assert np.array_equal(ism, mA) and list(iids[: len(idsA)]) == list(idsA)


@pytest.mark.gpu
def test_memory_admission_evicts_and_stays_exact():
    vols = _volumes()
    F = (DIMS[0] // DS) * (DIMS[1] // DS) * (DIMS[2] // DS)
    per = K * F * 4
    with _server(vols, slots=1) as s:
        ref = [s.wait(s.submit(*r))[0] for r in REQS]
    with _server(vols, slots=3, mem_budget=2 * per + 1) as s:
        got = [s.wait(t)[0] for t in [s.submit(*r) for r in REQS]]
        st = s.stats()
        assert st.evictions > 0 and st.cache_bytes <= 2 * per + per  # one in flight may exceed
    for a, b in zip(ref, got):
        assert np.array_equal(a, b)


@pytest.mark.gpu
def test_requests_report_what_they_computed():
    # a request counts the ISMs it computed and their map pairs; one answered
    # from the cache counts none (the server's field blocks come from its
    # pool, so repeated requests allocate nothing)
    vols = _volumes()
    with _server(vols, slots=1) as s:
        _, _, info = s.wait(s.submit(isoamd.REQ_ISM, 0, 0))
        m = info.m
        assert info.computed == 1 and info.computed_pairs == m * (m - 1) // 2
        _, _, info = s.wait(s.submit(isoamd.REQ_ISM, 0, 0))
        assert info.computed == 0 and info.computed_pairs == 0
        _, _, tsm = s.wait(s.submit(isoamd.REQ_TSM, 0, 0, 0, 1))  # (0, 1) is new
        assert tsm.computed == 1
        _, _, ism = s.wait(s.submit(isoamd.REQ_ISM, 0, 1))
        assert ism.computed == 0 and tsm.computed_pairs == ism.m * (ism.m - 1) // 2


@pytest.mark.gpu
def test_requests_never_measure_launch_shapes(ctx):
    # volumes of very different surface sizes fall into different autotuner
    # shape classes; the server measures launch shapes only while it starts,
    # so requests take the nearest measured choice, and every map equals the
    # one a context that measured its own choices computes
    vols = {(0, 0): isoamd.gen_timeseries(*DIMS, 0, 0, nblobs=1, seed=9),
            (0, 1): np.random.default_rng(3).random(DIMS).astype(np.float32),  # white noise
            (0, 2): isoamd.gen_timeseries(*DIMS, 0, 2, nblobs=40, seed=9)}
    with _server(vols, slots=1) as s:  # the start measures on the first volume only
        got = [s.wait(s.submit(isoamd.REQ_ISM, 0, t)) for t in range(3)]
        st = s.stats()
    assert st.tunings_in_requests == 0
    assert st.untuned_shapes > 0  # the noise volume's surfaces are far larger
    for t, (mp, ids, info) in enumerate(got):
        img = vols[(0, t)]
        lo, hi = float(img.min()), float(img.max())
        isos = np.array([lo + (i + 1) * (np.float64(hi) - lo) / (K + 1) for i in range(K)],
                        np.float32)
        smap, want_ids, stt, _ = ctx.ism_volume(ctx.to_device(img), DIMS, isos, isoamd.EXTRACT_MC,
                                                DS, H)
        m = stt.n_nonempty
        assert info.m == m
        assert np.array_equal(mp, smap.numpy()[: m * m].reshape(m, m))
        assert list(ids[:m]) == list(want_ids[:m])
