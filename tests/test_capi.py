"""The C-ABI library loads and exports every symbol include/iso_amd.h declares
(no GPU needed: nothing is called that touches a device)."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def _declared():
    src = open(os.path.join(ROOT, "include", "iso_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(iso_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = _declared()
    for must in ("iso_mc_extract", "iso_approx_extract", "iso_df_points_f3", "iso_hist_nmi"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import isoamd

    L = isoamd.lib()
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, missing
    assert set(isoamd.EXPORTS) <= set(_declared())


def test_status_strings_without_device():
    import isoamd

    L = isoamd.lib()
    assert L.iso_status_string(0) == b"ok"
    assert b"capacity" in L.iso_status_string(4)


def test_generators_are_deterministic():
    import isoamd

    a = isoamd.gen_blobs(16, 12, 10, seed=7)
    b = isoamd.gen_blobs(16, 12, 10, seed=7)
    c = isoamd.gen_blobs(16, 12, 10, seed=8)
    assert a.shape == (16, 12, 10) and np.array_equal(a, b) and not np.array_equal(a, c)
    f = isoamd.gen_front(40, 20, 20, seed=2)
    assert f.min() >= 0.0 and f.max() <= 1.0
    # tanh front: low on the x=0 side, high on the far side
    assert f[:5].mean() < 0.2 and f[-5:].mean() > 0.8
    s = isoamd.gen_sines(16, 16, 16, seed=4)
    assert np.abs(s).max() <= 4.0


@pytest.mark.gpu
def test_device_copy_aligned_and_ragged(ctx):
    rng = np.random.default_rng(3)
    a = rng.random(4099).astype(np.float32)
    src = ctx.to_device(a)
    dst = ctx.empty(a.shape, np.float32)
    ctx.copy(dst, src, 4096 * 4)  # kernel path
    assert np.array_equal(dst.numpy()[:4096], a[:4096])
    ctx.copy(dst, src)  # 4099 floats: not a multiple of 16 bytes
    assert np.array_equal(dst.numpy(), a)


def test_reference_host_library_links_against_the_mirror(tmp_path):
    # every symbol the reference's isosurfaces_lib takes from its CUDA library
    # (funcs.cpp's five extern "C" calls and Isosurface.cpp's cudaMarchingCube)
    # resolves in libisosurfaces_compat.so: an executable referencing all of
    # them links with no undefined symbol
    import subprocess

    pkg = os.path.join(ROOT, "isosurface-analysis_amd")
    if not os.path.exists(os.path.join(pkg, "libisosurfaces_compat.so")):
        pytest.skip("libisosurfaces_compat.so is not built")
    exe = tmp_path / "link_check"
    r = subprocess.run(["/opt/rocm/bin/hipcc", "-O0", "-std=c++17",
                        os.path.join(ROOT, "tests", "link_reference_symbols.cpp"), "-o", str(exe),
                        "-L" + pkg, "-lisosurfaces_compat", "-liso_amd",
                        "-Wl,-rpath," + pkg, "-Wl,--no-undefined", "-Wl,--no-allow-shlib-undefined"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    nm = subprocess.run(["nm", "-D", "--defined-only",
                         os.path.join(pkg, "libisosurfaces_compat.so")],
                        capture_output=True, text=True, check=True).stdout
    for sym in ("calculate_distances_h", "approximate_isosurfaces_h", "calculate_histogram_h",
                "generateHierarchy_h", "findExtrema_h"):
        assert re.search(r"\bT " + sym + r"\b", nm), sym
