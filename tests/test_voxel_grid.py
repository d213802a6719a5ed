"""Voxel-grid downsampling (pcl::VoxelGrid as the reference driver calls it,
src/alignment.cpp:78-87,259-288).  The oracle restatement is pinned against
an independent brute force (voxel of each point from floor(p / leaf) in
float64 ... checked to agree with the float32 index where the two can
differ only at exact voxel boundaries, means in float64) and special cases;
the device path must equal the oracle bit for bit."""
import numpy as np
import pytest

import ppf_oracle as O


def _brute(points, normals, leaf):
    p = np.asarray(points, np.float32)
    n = np.asarray(normals, np.float32)
    ok = np.isfinite(p).all(axis=1)
    p, n = p[ok], n[ok]
    if len(p) == 0:
        return np.zeros((0, 3)), np.zeros((0, 3))
    inv = np.float32(1.0) / np.float32(leaf)
    ijk = np.floor(p * inv).astype(np.int64)          # float32 products, as PCL
    lo = ijk.min(axis=0)
    dims = ijk.max(axis=0) - lo + 1
    idx = (ijk[:, 0] - lo[0]) + (ijk[:, 1] - lo[1]) * dims[0] + (ijk[:, 2] - lo[2]) * dims[0] * dims[1]
    groups = {}
    for k, i in enumerate(idx.tolist()):
        groups.setdefault(i, []).append(k)
    keys = sorted(groups)
    mp = np.array([p[groups[k]].astype(np.float64).mean(axis=0) for k in keys])
    mn = np.array([n[groups[k]].astype(np.float64).mean(axis=0) for k in keys])
    return mp, mn


@pytest.mark.parametrize("seed,n,leaf", [(0, 2000, 0.1), (1, 5000, 0.03), (2, 300, 1.5), (3, 4000, 0.25)])
def test_oracle_matches_brute_force(seed, n, leaf):
    rng = np.random.default_rng(seed)
    p = (rng.normal(size=(n, 3)) * 2.0).astype(np.float32)
    nn = rng.normal(size=(n, 3)).astype(np.float32)
    # duplicates and points exactly on voxel boundaries
    p[:50] = p[50:100]
    p[100:120] = (np.round(p[100:120] / leaf) * leaf).astype(np.float32)
    op, on = O.voxel_downsample(p, nn, leaf)
    bp, bn = _brute(p, nn, leaf)
    assert op.shape == bp.shape
    np.testing.assert_allclose(op, bp, rtol=1e-5, atol=1e-5 * leaf)
    np.testing.assert_allclose(on, bn, rtol=1e-4, atol=1e-5)


def test_oracle_special_cases():
    p = np.array([[0.1, 0.1, 0.1], [0.12, 0.13, 0.11], [np.nan, 0, 0], [5, 5, 5]], np.float32)
    n = np.array([[0, 0, 1], [0, 1, 0], [1, 0, 0], [1, 0, 0]], np.float32)
    op, on = O.voxel_downsample(p, n, 1.0)
    # NaN point dropped; first two share voxel (0,0,0); normals averaged, not renormalized
    assert len(op) == 2
    np.testing.assert_array_equal(on[0], np.float32([0, 0.5, 0.5]))
    np.testing.assert_array_equal(op[1], p[3])
    # a single point and an empty cloud
    op, _ = O.voxel_downsample(p[:1], n[:1], 0.5)
    np.testing.assert_array_equal(op, p[:1])
    op, _ = O.voxel_downsample(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32), 0.5)
    assert len(op) == 0
    # leaf so small that the box holds more than INT32_MAX voxels: input as is
    q = np.array([[0, 0, 0], [1000, 1000, 1000]], np.float32)
    op, on = O.voxel_downsample(q, q, 1e-3)
    np.testing.assert_array_equal(op, q)


@pytest.mark.gpu
@pytest.mark.parametrize("seed,n,leaf", [(0, 2000, 0.1), (5, 200000, 0.02), (6, 50000, 0.3)])
def test_device_matches_oracle(ctx, seed, n, leaf):
    import ppf_amd
    rng = np.random.default_rng(seed)
    p = (rng.normal(size=(n, 3)) * 2.0).astype(np.float32)
    nn = rng.normal(size=(n, 3)).astype(np.float32)
    p[::97, 1] = np.nan
    op, on = ppf_amd.voxel_downsample((p, nn), leaf, ctx=ctx)
    ep, en = O.voxel_downsample(p, nn, leaf)
    assert op.tobytes() == ep.tobytes() and on.tobytes() == en.tobytes()


@pytest.mark.gpu
def test_device_special_cases(ctx):
    import ppf_amd
    q = np.array([[0, 0, 0], [1000, 1000, 1000]], np.float32)
    op, _ = ppf_amd.voxel_downsample((q, q), 1e-3, ctx=ctx)
    np.testing.assert_array_equal(op, q)
    nanc = np.full((4, 3), np.nan, np.float32)
    op, _ = ppf_amd.voxel_downsample((nanc, nanc), 0.1, ctx=ctx)
    assert len(op) == 0


def test_bbox_extent():
    import ppf_amd
    p = np.array([[0, 1, 2], [3, -1, 2.5], [np.inf, 0, 0]], np.float32)
    np.testing.assert_array_equal(ppf_amd.bbox_extent(p), np.float32([3, 2, 0.5]))
