"""The `distance` module's isovalue equalisation restated over the oracle's
stages -- TEST INFRASTRUCTURE ONLY (imported by tests/ as the checker).

Follows runtime/src/DistanceEqualization.cpp with the switches of
runtime/hdr/DistanceEqualization.h:19-53: run() (:626-773) evaluates the
start table (.h:221, replacing the computed isovalues at .cpp:678-681), then
binarySearchEqualization() (:1451-1605) refines with binaryIsoValueSearch()
(:1338-1449); approximation() (:84-186) walks empty isosurfaces' isovalues;
computeDistances() (:560-624) takes d_j = 1 - NMI(j, j+1) of the 17-field,
17-bin map; evaluateDistances() (:333-558) keeps the best moving values.
Every float operation is a float32 operation, as in the reference.
"""
from __future__ import annotations

import numpy as np

from . import approx, df_tree, ism, normalize

F32 = np.float32
NUM_DIST = 16
START = np.array([137.58087158203125, 137.7071075439453125, 138.2776336669921875,
                  140.0638885498046875, 144.370941162109375, 151.1020965576171875,
                  157.6292877197265625, 163.8123931884765625, 174.6049346923828125,
                  214.08880615234375, 324.660186767578125, 604.75189208984375,
                  1200.432861328125, 2765.340087890625, 6707.39599609375, 14770.724609375,
                  33156.9921875], np.float32)


def intersect(lb, ub):
    """binaryIsoValueSearch's new isovalue (DistanceEqualization.cpp:1421-1433):
    where the line of left distances and the line of right distances through
    the lower bound (value, left, right) and the upper bound cross."""
    with np.errstate(all="ignore"):
        i1 = F32(F32(lb[1] - ub[1]) / F32(lb[0] - ub[0]))
        f1 = F32(lb[1] - F32(i1 * lb[0]))
        i2 = F32(F32(lb[2] - ub[2]) / F32(lb[0] - ub[0]))
        f2 = F32(lb[2] - F32(i2 * lb[0]))
        return F32(F32(f2 - f1) / F32(i1 - i2))


def run(img: np.ndarray, ds: int, max_itr: int = 180, distances_fn=None, cells_fn=None):
    """img: the z-fastest image (X, Y, Z).  Returns the isovalue sets and the
    distances of the estimation step and of each refinement step.

    distances_fn(iso) -> 16 distances and cells_fn(isovalue) -> cell count
    replace the pipeline stages (voxel approximation, fields, the 17-bin map)
    when given; the loop itself is unchanged."""
    dims = img.shape
    rmin, rmax = F32(img.min()), F32(img.max())
    k = NUM_DIST + 1
    fixed = [False] * k
    fixed[0] = fixed[NUM_DIST] = True
    pushed = [False] * k
    st = {"best_avg": F32(10e20), "best_max": F32(10e20), "best_iso": START.copy(),
          "moving": [[F32(0), F32(0)] for _ in range(NUM_DIST)], "it": 0}
    dist = np.zeros(NUM_DIST, np.float32)

    def cells(v):
        if cells_fn is not None:
            return cells_fn(F32(v))
        return approx(img, F32(v)).shape[0]

    def approximation(iso):
        for d in range(k):
            cur = F32(iso[d])
            n = cells(cur)
            val = F32(iso[d + 1]) if d < NUM_DIST else F32(cur + F32(rmax - rmin) / F32(NUM_DIST))
            incr = F32(F32(val - cur) / F32(NUM_DIST))
            while n == 0:
                cur = F32(cur + incr)
                if cur > rmax or cur != cur:
                    assert d > 0
                    cur = F32(iso[d - 1])
                    incr = F32(F32(iso[d] - cur) / F32(16))
                    cur = F32(cur + incr)
                if d < NUM_DIST and cur > iso[d + 1]:
                    iso[d] = iso[d + 1]
                    iso[d + 1] = cur
                    cur = F32(iso[d])
                iso[d] = cur
                n = cells(cur)

    def distances(iso):
        if distances_fn is not None:
            dist[:] = np.asarray(distances_fn(iso.copy()), np.float32)
            return
        fields = np.stack([df_tree(normalize(approx(img, F32(v)), dims), dims, ds) for v in iso])
        m32, _ = ism(fields, k)  # the reference's float-order NMI (JointHistogram.cu:24-60)
        for j in range(NUM_DIST):
            dist[j] = F32(1) - m32[j, j + 1]

    def evaluate(iso, fix_even):
        acc = F32(0)
        for j in range(NUM_DIST):
            acc = F32(acc + dist[j])
        avg = F32(acc / F32(NUM_DIST))
        acc = F32(0)
        maxe = F32(-1)
        for j in range(NUM_DIST):
            e = F32(abs(F32(dist[j] - avg)))
            acc = F32(acc + e)
            maxe = max(maxe, e)
        avge = F32(acc / F32(NUM_DIST))
        if F32(avge / avg) < st["best_avg"]:
            st["best_avg"] = F32(avge / avg)
            st["best_iso"] = iso.copy()
        if maxe < st["best_max"]:
            st["best_max"] = maxe
        mv = st["moving"]
        for i in range(1 if fix_even else 0, NUM_DIST - 1, 2):
            if i > 0:
                dl, dr = dist[i - 1], dist[i]
                if F32(abs(F32(dl - dr))) < F32(abs(F32(mv[i - 1][1] - mv[i][1]))):
                    mv[i - 1] = [F32(iso[i - 1]), dl]
                    mv[i] = [F32(iso[i]), dr]

    def fix_inversions(iso):
        for i in range(k):
            for j in range(i + 1, k):
                if iso[i] > iso[j]:
                    iso[i], iso[j] = iso[j], iso[i]
                    fixed[i], fixed[j] = fixed[j], fixed[i]
                    pushed[i], pushed[j] = pushed[j], pushed[i]

    infos = []

    def search(fix_even, iso, info):
        info_no = st["it"] % 5
        for i in range(1 if fix_even else 2, NUM_DIST, 2):
            j, kk = i - 1, i + 1
            dji, dik = dist[j], dist[i]
            s = F32(dji + dik)
            if info_no == 0:
                ub = (F32(iso[kk]), s, F32(0))
                lb = (F32(iso[j]), F32(0), s)
            else:
                ub = infos[st["it"] - 1]["upper"][i]
                lb = infos[st["it"] - 1]["lower"][i]
            if dji > dik:
                ub = (F32(iso[i]), dji, dik)
            if dji < dik:
                lb = (F32(iso[i]), dji, dik)
            nv = intersect(lb, ub)
            if not np.isnan(nv) and not fixed[i]:
                iso[i] = nv
            info["upper"][i] = ub
            info["lower"][i] = lb

    out = []
    st["it"] = 1
    iso = START.copy()
    approximation(iso)
    distances(iso)
    evaluate(iso, False)
    out.append(("final", iso.copy(), dist.copy()))
    st["it"] = 0
    fix_even = True
    while True:
        if st["it"] == 0:
            fix_even = True
            iso = st["best_iso"].copy()
            st["moving"] = [[F32(st["best_iso"][j]), F32(dist[j])] for j in range(NUM_DIST)]
        else:
            iso = infos[-1]["iso"].copy()
        info = {"upper": {}, "lower": {}}
        fix_inversions(iso)
        search(fix_even, iso, info)
        approximation(iso)
        distances(iso)
        finished = st["it"] == max_itr
        evaluate(iso, fix_even)
        info["iso"] = iso.copy()
        infos.append(info)
        out.append(("final" if finished else st["it"], iso.copy(), dist.copy()))
        if finished:
            return out
        st["it"] += 1
        if st["it"] % 5 == 0:
            for i in range(NUM_DIST):
                dist[i] = st["moving"][i][1]
            fix_even = not fix_even
