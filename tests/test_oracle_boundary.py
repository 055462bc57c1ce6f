"""Pins the oracle's restatements of the two remaining entry points of the
reference's CUDA library that its host code links against:
calculate_distances_h (isosurfaces_cuda/src/DistanceField.cu:15-98) and
findExtrema_h (src/IsosurfaceExtrema.cu:22-138), with closed forms."""
import numpy as np

import oracle


def _samples(grid, ds):
    nx, ny, nz = grid
    x, y, z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    return np.stack([x * ds + ds // 2, y * ds + ds // 2, z * ds + ds // 2], -1).reshape(-1, 3)


def test_voxel_units_single_point_is_the_euclidean_norm():
    # integer coordinates below 2^12: every square and sum is exact, so the
    # field is the correctly rounded root of the exact squared distance
    grid, ds = (6, 5, 4), 2
    p = np.array([[3.0, 1.0, 2.0]], np.float32)
    got = oracle.df_voxel_units(p, grid, ds)
    d2 = ((_samples(grid, ds) - p) ** 2).sum(1)
    assert np.array_equal(got, np.sqrt(d2.astype(np.float64)).astype(np.float32))


def test_voxel_units_plane_and_empty_set():
    grid, ds = (7, 4, 3), 1
    ys, zs = np.meshgrid(np.arange(4), np.arange(3), indexing="ij")
    plane = np.stack([np.full(12, 2.0), ys.ravel(), zs.ravel()], 1).astype(np.float32)
    got = oracle.df_voxel_units(plane, grid, ds)
    # ds=1: samples at x + 0 (ds/2 = 0 in integer arithmetic)
    want = np.abs(_samples(grid, ds)[:, 0] - 2.0).astype(np.float32)
    assert np.array_equal(got, want)
    none = oracle.df_voxel_units(np.zeros((0, 3), np.float32), grid, ds)
    assert np.all(none == np.sqrt(np.float32(1e38)))


def test_voxel_units_point_runs_keep_the_minimum():
    # 2,000,001 points: the last one, alone in the second run, is the nearest
    # to every sample, so each field value is its distance
    grid, ds = (3, 3, 2), 2
    n = 2_000_001
    rng = np.random.default_rng(2)
    pts = (rng.random((n, 3)) * 50 + 100).astype(np.float32)
    pts[-1] = (2.0, 3.0, 1.0)
    got = oracle.df_voxel_units(pts, grid, ds)
    d2 = ((_samples(grid, ds) - pts[-1]) ** 2).sum(1)
    assert np.array_equal(got, np.sqrt(d2.astype(np.float64)).astype(np.float32))


def test_extrema_of_a_ramp_are_its_two_corners():
    # f = x + 10 y + 100 z (cube, the reference index is x-fastest): every
    # voxel but the two extreme corners sees a lower and a higher neighbour
    n = 6
    x, y, z = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    f = np.zeros(n ** 3, np.float32)
    f[(x + y * n + z * n * n).ravel()] = (x + 10 * y + 100 * z).ravel()
    assert list(oracle.find_extrema(f, (n, n, n))) == [0.0, 555.0]


def test_extrema_of_constant_volume_and_plateau_rule():
    assert oracle.find_extrema(np.full(125, 3.0, np.float32), (5, 5, 5)).size == 0
    # a single high voxel in a low volume: a strict maximum (one equal, itself)
    f = np.zeros(216, np.float32)
    f[2 + 2 * 6 + 2 * 36] = 7.0
    got = oracle.find_extrema(f, (6, 6, 6))
    assert list(got) == [7.0]  # every zero voxel sees more than 5 equal zeros
