"""Runtime modules (isosurfaces_amd single / script / multi / distance): output
layout, run-information file, resume skipping, and on the GPU the written map,
the representative fields, the cross maps and the distance equalisation's
steps against the oracle pipeline."""
import json
import os
import subprocess

import numpy as np
import pytest

import isoamd
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "isosurface-analysis_amd", "isosurfaces_amd")
NMI_RTOL, NMI_ATOL = 1e-6, 1e-7


def _run(module, cfg_path, expect=0):
    if not os.path.exists(BIN):
        pytest.skip("isosurfaces_amd not built")
    r = subprocess.run([BIN, module, str(cfg_path)], capture_output=True, text=True, timeout=300)
    assert r.returncode == expect, r.stdout + r.stderr
    return r


def _write_json(path, obj):
    with open(path, "w") as f:
        json.dump(obj, f)
    return path


def _run_info(tmp, variables, volumes, maps, dims=(40, 32, 24), ds=2):
    return {"dimensions": list(dims), "variables": variables, "dfDownscale": ds,
            "volumes": volumes, "similarityMaps": maps}


# ---- host-only paths (no device work happens on them) ----------------------

def test_single_skips_processed_volume(tmp_path):
    out = tmp_path / "out"
    out.mkdir()
    meta = _run_info(tmp_path, ["GT"], [{"variable": 0, "timestep": 150, "numIsovalues": 3,
                                         "representatives": [], "runtime": 1.0}], [])
    _write_json(out / "meta.json", meta)
    cfg = _write_json(tmp_path / "single.json", {
        "fileName": str(tmp_path / "missing"), "dimensions": [40, 32, 24], "minValue": 0.0,
        "maxValue": 1.0, "outputFolder": str(out), "variableName": "GT", "timestep": 150,
        "jsonFile": "meta.json", "dfDownscale": 2})
    r = _run("single", cfg)
    assert "already processed" in r.stdout
    assert json.load(open(out / "meta.json")) == meta  # untouched


def test_single_reports_short_file(tmp_path):
    out = tmp_path / "out"
    (tmp_path / "v.dat").write_bytes(np.zeros(10, np.float32).tobytes())
    cfg = _write_json(tmp_path / "single.json", {
        "fileName": str(tmp_path / "v"), "dimensions": [40, 32, 24], "minValue": 0.0,
        "maxValue": 1.0, "outputFolder": str(out), "variableName": "GT", "timestep": 1,
        "jsonFile": "meta.json", "dfDownscale": 2})
    r = _run("single", cfg, expect=2)
    assert "bytes, need" in r.stderr


def test_script_skips_missing_volumes_and_checks_format(tmp_path):
    entry = {"dimensions": [40, 32, 24],
             "variables": [{"name": "a", "minValue": 0, "maxValue": 1, "variableFolder": "fa"}],
             "timesteps": [1, 2], "dfDownscale": 2, "fileFormatString": "%s/%s/%s-%04d",
             "inputFolder": str(tmp_path), "outputFolder": str(tmp_path / "out"),
             "jsonFile": "meta.json"}
    r = _run("script", _write_json(tmp_path / "s.json", {"entries": [entry]}))
    assert f"{tmp_path}/fa/a-0001" in r.stdout and f"{tmp_path}/fa/a-0002" in r.stdout
    entry["fileFormatString"] = "%d/%s/%s/%s"
    r = _run("script", _write_json(tmp_path / "s2.json", {"entries": [entry]}), expect=2)
    assert "fileFormatString" in r.stderr


def test_multi_without_fields_adds_nothing(tmp_path):
    root = tmp_path / "data"
    root.mkdir()
    meta = _run_info(tmp_path, ["a", "b"], [], [])
    _write_json(root / "meta.json", meta)
    cfg = _write_json(tmp_path / "m.json", {"jsonFile": "meta.json", "timeSteps": [0, 1],
                                            "dataRoot": str(root), "dfDownscale": 2})
    r = _run("multi", cfg)
    assert r.stdout.count("failed") == 4  # VSM a-b at t0, t1; TSM a and b over (0, 1)
    assert json.load(open(root / "meta.json"))["similarityMaps"] == []
    assert os.path.isdir(str(root) + "-multi-tmp/a-0")


# ---- full runs on the GPU ----------------------------------------------------

DIMS = (40, 32, 24)
DS = 2


def _volume(var, step):
    return isoamd.gen_timeseries(*DIMS, var, step, nblobs=5, seed=17)


def _write_volume(path, img):
    # files are x-fastest (BinaryFloatReader.cpp:29-49)
    np.ascontiguousarray(img.transpose(2, 1, 0)).tofile(path)


def _oracle_volume(img, lo_cfg, hi_cfg, approx):
    X, Y, Z = DIMS
    file = np.ascontiguousarray(img.transpose(2, 1, 0)).ravel()
    zimg, rmin, rmax = oracle.ingest(file, DIMS, lo_cfg, hi_cfg)
    isos = oracle.equalized_isovalues(zimg, rmin, rmax, 256)
    ids, fields = [], []
    for i, iso in enumerate(isos):
        if approx:
            pts = oracle.approx(zimg, iso)
        else:
            pts, _ = oracle.mc(zimg.ravel(), DIMS, iso, ncls=oracle.mc_ncls_compat(X, Y, Z))
        if pts.shape[0]:
            ids.append(i)
            fields.append(oracle.df_tree(oracle.normalize(pts, DIMS), DIMS, DS))
    return isos, ids, np.stack(fields)


def _check_volume(out, var, t, approx, img, lo_cfg, hi_cfg):
    app = "256--approx-scale_1" if approx else "256-"
    vdir = os.path.join(out, f"{var}-{t}")
    isos, ids, fields = _oracle_volume(img, lo_cfg, hi_cfg, approx)
    m = len(ids)
    got = np.fromfile(os.path.join(vdir, f"similarityMap-{app}", "map.dat"), np.float32)
    assert got.size == m * m
    got = got.reshape(m, m)
    m32, _ = oracle.ism(fields, 128)  # the reference's float order, bit for bit
    assert np.array_equal(got, m32)
    meta = json.load(open(os.path.join(out, "meta.json")))
    vi = meta["variables"].index(var)
    vol = [v for v in meta["volumes"] if v["variable"] == vi and v["timestep"] == t]
    assert len(vol) == 1 and vol[0]["numIsovalues"] == m
    reps = vol[0]["representatives"]
    want = oracle.representatives(got, 16, ids)
    assert [r["valueId"] for r in reps] == [w[0] for w in want]
    assert [r["mapId"] for r in reps] == [w[1] for w in want]
    assert [r["repId"] for r in reps] == list(range(len(want)))
    np.testing.assert_array_equal(np.float32([r["importance"] for r in reps]),
                                  np.float32([w[2] for w in want]))
    np.testing.assert_array_equal(np.float32([r["isovalue"] for r in reps]),
                                  isos[[w[0] for w in want]])
    for r in reps:
        assert r["filename"].startswith(os.path.join(vdir, f"isosurfaces-{app}", ""))
    asc = sorted(w[0] for w in want)
    F = (DIMS[0] // DS) * (DIMS[1] // DS) * (DIMS[2] // DS)
    for rank, vid in enumerate(asc):
        df = np.fromfile(os.path.join(vdir, "distancefields", f"distancefield-{rank}.df"), np.float32)
        assert df.size == F and np.array_equal(df, fields[ids.index(vid)])
    assert {"var1": vi, "t1": t, "var2": vi, "t2": t,
            "filename": os.path.join(vdir, f"similarityMap-{app}", "map.dat")} in meta["similarityMaps"]


@pytest.mark.gpu
def test_single_module_matches_oracle(tmp_path):
    img = _volume(0, 3)
    _write_volume(tmp_path / "v0.dat", img)
    lo_cfg, hi_cfg = float(img.min()), float(img.max())
    out = str(tmp_path / "out")
    cfg = _write_json(tmp_path / "single.json", {
        "fileName": str(tmp_path / "v0"), "dimensions": list(DIMS), "minValue": lo_cfg,
        "maxValue": hi_cfg, "outputFolder": out, "variableName": "f", "timestep": 3,
        "jsonFile": "meta.json", "dfDownscale": DS})
    print("module single", flush=True)
    _run("single", cfg)
    _check_volume(out, "f", 3, True, img, lo_cfg, hi_cfg)
    r = _run("single", cfg)  # second run: resume skip
    assert "already processed" in r.stdout


@pytest.mark.gpu
def test_script_then_multi_match_oracle(tmp_path):
    data = tmp_path / "in"
    out = tmp_path / "out"
    variables = []
    imgs = {}
    for v, name in enumerate(("f", "g")):
        (data / name).mkdir(parents=True)
        for t in (0, 1):
            imgs[(name, t)] = _volume(v, t)
            _write_volume(data / name / f"{name}{t:02d}.dat", imgs[(name, t)])
        lo = min(float(imgs[(name, t)].min()) for t in (0, 1))
        hi = max(float(imgs[(name, t)].max()) for t in (0, 1))
        variables.append({"name": name, "minValue": lo, "maxValue": hi, "variableFolder": name})
    entry = {"dimensions": list(DIMS), "variables": variables, "timesteps": [0, 1],
             "dfDownscale": DS, "fileFormatString": "%s/%s/%s%02d", "inputFolder": str(data),
             "outputFolder": str(out), "jsonFile": "meta.json"}
    print("module script", flush=True)
    _run("script", _write_json(tmp_path / "script.json", {"entries": [entry]}))
    for var in variables:
        for t in (0, 1):
            _check_volume(str(out), var["name"], t, False, imgs[(var["name"], t)],
                          var["minValue"], var["maxValue"])
    print("module multi", flush=True)
    cfg = _write_json(tmp_path / "multi.json", {"jsonFile": "meta.json", "timeSteps": [0, 1],
                                                "dataRoot": str(out), "dfDownscale": DS})
    _run("multi", cfg)
    meta = json.load(open(out / "meta.json"))
    F = (DIMS[0] // DS) * (DIMS[1] // DS) * (DIMS[2] // DS)

    def reps(name, t):
        d = out / f"{name}-{t}" / "distancefields"
        return np.stack([np.fromfile(d / f"distancefield-{i}.df", np.float32) for i in range(16)])

    cases = [("f", 0, "g", 0), ("f", 1, "g", 1), ("f", 0, "f", 1), ("g", 0, "g", 1)]
    for v1, t1, v2, t2 in cases:
        fn = f"{out}-multi-tmp/{v1}-{t1}/TMP-multiMap-16-{v1}-{t1}-{v2}-{t2}.dat"
        got = np.fromfile(fn, np.float32).reshape(16, 16)
        m32, _ = oracle.ism(np.concatenate([reps(v1, t1), reps(v2, t2)]), 128, multi=True)
        assert np.array_equal(got, m32)
        entry = {"var1": meta["variables"].index(v1), "t1": t1,
                 "var2": meta["variables"].index(v2), "t2": t2, "filename": fn}
        assert entry in meta["similarityMaps"]
    assert F == 20 * 16 * 12
    r = _run("multi", cfg)  # everything present: nothing recomputed
    assert r.stdout.strip() == ""


# ---- distance (DistanceEqualization) ---------------------------------------

def _distance_setup(tmp_path, dims=(48, 40, 32), ds=2, ts=3):
    # a front volume over the value range the module's start table is made
    # for (runtime/hdr/DistanceEqualization.h:221, 137.6 .. 33157)
    f = isoamd.gen_front(*dims, seed=4)
    img = (np.float32(72.16) + f.astype(np.float32) * np.float32(37270 - 72.16)).astype(np.float32)
    (tmp_path / "gt.dat").write_bytes(np.ascontiguousarray(img.transpose(2, 1, 0)).tobytes())
    out = str(tmp_path / "out") + "/"
    cfg = _write_json(tmp_path / "d.json", {
        "fileName": str(tmp_path / "gt"), "dimensions": list(dims), "minValue": 72.16,
        "maxValue": 37270.0, "outputFolder": out, "variableName": "GT", "timestep": ts,
        "jsonFile": "meta.json", "dfDownscale": ds})
    return img, cfg, out


def _csv_rows(path):
    return [l.rstrip("\n").split(",") for l in open(path)]


def test_distance_skips_processed_timestep(tmp_path):
    out = tmp_path / "out"
    out.mkdir()
    # any variable of the time step counts (DistanceEqualization.cpp:971-995)
    meta = _run_info(tmp_path, ["other"], [{"variable": 0, "timestep": 3, "numIsovalues": 17,
                                            "representatives": [], "runtime": 0.0}], [])
    _write_json(out / "meta.json", meta)
    cfg = _write_json(tmp_path / "d.json", {
        "fileName": str(tmp_path / "missing"), "dimensions": [40, 32, 24], "minValue": 0.0,
        "maxValue": 1.0, "outputFolder": str(out) + "/", "variableName": "GT", "timestep": 3,
        "jsonFile": "meta.json", "dfDownscale": 2})
    r = _run("distance", cfg)
    assert "already processed" in r.stdout
    assert json.load(open(out / "meta.json")) == meta
    # the folders and CSV headers are made before the check, as in the reference
    head = _csv_rows(out / "GT-3" / "isovalues.csv")
    assert head == [["iterations"] + [f"iv{i}" for i in range(17)] + [""]]


@pytest.mark.gpu
def test_distance_module_follows_oracle(tmp_path):
    # the first 12 refinement steps against the oracle restatement: the same
    # isovalues and distances at every step, bit for bit (exact fields and
    # counts, the reference's float-order NMI on both sides, the loop's float
    # arithmetic in the same order; the CSV holds 32 significant digits)
    from oracle import distance

    img, cfg, out = _distance_setup(tmp_path)
    env = dict(os.environ, ISO_DISTANCE_MAX_ITR="12")
    r = subprocess.run([BIN, "distance", str(cfg)], capture_output=True, text=True, timeout=300,
                       env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    want = distance.run(img, 2, max_itr=12)
    rows = _csv_rows(out + "GT-3/isovalues.csv")[1:]
    drows = [r_ for r_ in _csv_rows(out + "GT-3/distances.csv")[1:]]
    assert len(rows) == len(want) == 14 and len(drows) == 14
    for (tag, iso, dist), row, drow in zip(want, rows, drows):
        assert row[0] == str(tag)
        got = np.array([float(v) for v in row[1:18]], np.float32)
        np.testing.assert_array_equal(got, iso)
        gd = np.array([float(v) for v in drow[1:17]], np.float32)
        np.testing.assert_array_equal(gd, dist)


@pytest.mark.gpu
def test_distance_module_full_run_layout(tmp_path):
    # the reference's 181 refinement steps: files, line counts, run entries
    img, cfg, out = _distance_setup(tmp_path, ts=7)
    _run("distance", cfg)
    base = out + "GT-7/"
    for d in ("isosurfaces-17--approx-scale_1", "similarityMap-17--approx-scale_1",
              "distancefields"):
        assert os.path.isdir(base + d)
    iso = _csv_rows(base + "isovalues.csv")
    assert len(iso) == 1 + 1 + 181 and iso[1][0] == "final" and iso[-1][0] == "final"
    assert [r_[0] for r_ in iso[2:-1]] == [str(i) for i in range(180)]
    assert len(_csv_rows(base + "distances.csv")) == 183
    jumps = _csv_rows(base + "jumps.csv")
    assert len(jumps) == 183 and jumps[1][:2] == ["0", "iterations"]
    assert len(_csv_rows(base + "distances-split.csv")) == 1 + 2 * 182
    meta = json.load(open(out + "meta.json"))
    assert meta["variables"] == ["GT-0.660000"]
    assert len(meta["volumes"]) == 20  # every 10th step and the final one
    last = meta["volumes"][-1]["representatives"]
    assert [r_["filename"] for r_ in last] == [f"180-surface-{i}.dat" for i in range(17)]
    assert [np.float32(r_["isovalue"]) for r_ in last] == [np.float32(v) for v in iso[-1][1:18]]
