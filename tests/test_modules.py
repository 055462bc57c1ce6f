"""Runtime modules (isosurfaces_amd single / script / multi): output layout,
run-information file, resume skipping, and on the GPU the written map, the
representative fields and the cross maps against the oracle pipeline."""
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
    _, m64 = oracle.ism(fields, 128)
    np.testing.assert_allclose(got, m64, rtol=NMI_RTOL, atol=NMI_ATOL)
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
        _, m64 = oracle.ism(np.concatenate([reps(v1, t1), reps(v2, t2)]), 128, multi=True)
        np.testing.assert_allclose(got, m64, rtol=NMI_RTOL, atol=NMI_ATOL)
        entry = {"var1": meta["variables"].index(v1), "t1": t1,
                 "var2": meta["variables"].index(v2), "t2": t2, "filename": fn}
        assert entry in meta["similarityMaps"]
    assert F == 20 * 16 * 12
    r = _run("multi", cfg)  # everything present: nothing recomputed
    assert r.stdout.strip() == ""
