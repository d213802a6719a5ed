"""GPU parity of the two rarely taken kernel paths, with device counters
showing that each one ran:

* k_vote's hit-by-hit path, taken by a reference row whose hits span more
  model buckets than the kernel's LDS grouping table holds (4,096), forced on
  small rows by lowering the grouping limit (ppf_debug_set knob 0);
* k_cluster_tiled's shared-hash path, taken by a chunk of poses that lie in
  several cells with one spatial hash (an FNV collision), on poses placed in
  two colliding cells (tests/collide.py).

Both are compared bit for bit with the CPU oracle."""
import numpy as np
import pytest

import collide
import ppf_amd
import ppf_oracle as O
import synth

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("use_l1,use_avg", [(False, False), (True, False), (False, True), (True, True)])
def test_shared_hash_chunks_match_oracle(ctx, use_l1, use_avg):
    T, codes, counts, d, _ = collide.colliding_poses(seed=3)
    print(f"shared-hash clustering: {len(T)} poses l1={use_l1} avg={use_avg}", flush=True)
    got = ctx.debug_cluster(T, codes, counts, d, use_l1=use_l1, use_avg=use_avg)
    ref = O.cluster_spatial(T, codes, counts, d, use_l1=use_l1, use_avg=use_avg)
    assert got["mixed_chunks"] > 0
    assert np.array_equal(_bits(got["vote_counts"]), _bits(ref["vote_counts"]))
    assert np.array_equal(_bits(got["trans"]), _bits(ref["trans"]))
    assert np.array_equal(_bits(got["quats"]), _bits(ref["quats"]))
    assert got["max_idx"] == ref["max_idx"]


def test_plain_poses_take_no_shared_hash_path(ctx):
    rng = np.random.default_rng(4)
    P = 500
    T = np.tile(np.eye(4, dtype=np.float32), (P, 1, 1))
    T[:, :3, 3] = rng.uniform(0, 3, (P, 3))
    codes = (np.arange(P, dtype=np.uint64) % 40) << np.uint64(6)
    counts = rng.integers(1, 9, P).astype(np.uint32)
    got = ctx.debug_cluster(T, codes, counts, 0.5)
    ref = O.cluster_spatial(T, codes, counts, 0.5)
    assert got["mixed_chunks"] == 0
    assert np.array_equal(_bits(got["vote_counts"]), _bits(ref["vote_counts"]))


def _quat_to_mat(q):
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], 1),
        np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], 1),
        np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], 1)], 1)


@pytest.mark.parametrize("P,cells", [(20000, 20), (70000, 200)])
def test_kd_grouped_cells_match_oracle(ctx, P, cells):
    """Cells of hundreds to a thousand poses: the clustering regroups each
    cell's poses as k-d leaves (level sorts of three passes below 2^16 poses,
    four above), which changes no result: vote counts, translations and the
    maximum equal the oracle's bit for bit."""
    rng = np.random.default_rng(P)
    d = 0.5
    base = rng.normal(size=(6, 4))
    base /= np.linalg.norm(base, axis=1, keepdims=True)
    q = base[rng.integers(0, 6, P)] + rng.normal(scale=0.05, size=(P, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    T = np.tile(np.eye(4), (P, 1, 1))
    T[:, :3, :3] = _quat_to_mat(q)
    side = int(np.ceil(cells ** (1 / 3)))
    cell = rng.integers(0, side, (P, 3))
    T[:, :3, 3] = (cell + rng.uniform(0.05, 0.95, (P, 3))) * d
    T = T.astype(np.float32)
    codes = np.arange(P, dtype=np.uint64) << np.uint64(6)
    counts = rng.integers(1, 50, P).astype(np.uint32)
    print(f"k-d grouped clustering: {P} poses in about {side ** 3} cells", flush=True)
    got = ctx.debug_cluster(T, codes, counts, d)
    ref = O.cluster_spatial(T, codes, counts, d)
    assert np.array_equal(_bits(got["vote_counts"]), _bits(ref["vote_counts"]))
    assert np.array_equal(_bits(got["trans"]), _bits(ref["trans"]))
    assert got["max_idx"] == ref["max_idx"]


@pytest.mark.parametrize("per_ref", [False, True])
def test_hit_by_hit_rows_match_oracle(per_ref):
    w = synth.plain_small(seed=3, n_model=500, n_clutter=700)
    m, s, d = w.models[0], w.scenes[0], w.d_dists[0]
    c = ppf_amd.Context(0)
    try:
        model = ppf_amd.Model(m, d, 0.3, ctx=c)
        scene = ppf_amd.Scene(s, d, 5, ctx=c)
        base = model.ppf_lookup(scene, per_reference_argmax=per_ref)
        assert base.info["vote_rows_ungrouped"] == 0
        bc, bn = base.votes()
        c.debug_set(0, 8)  # rows with more than 8 distinct buckets: hit by hit
        res = model.ppf_lookup(scene, per_reference_argmax=per_ref)
        print(f"rows voted hit by hit: {res.info['vote_rows_ungrouped']} of {scene.numPoints() // 5}",
              flush=True)
        assert res.info["vote_rows_ungrouped"] > 0
        codes, counts = res.votes()
        oc, on, st = O.vote(m, s, d, 5, 0.3, per_ref=per_ref)
        assert np.array_equal(codes, oc) and np.array_equal(counts, on)
        assert np.array_equal(codes, bc) and np.array_equal(counts, bn)
        assert res.info["votes"] == st["votes"] == base.info["votes"]
        assert res.pose().tobytes() == base.pose().tobytes()
        c.debug_set(0, 0)
        again = model.ppf_lookup(scene, per_reference_argmax=per_ref)
        assert again.info["vote_rows_ungrouped"] == 0
    finally:
        c.close()
