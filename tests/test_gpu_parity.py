"""GPU parity: libppf_amd.so (gfx950 kernels via the C ABI) against the CPU
oracle on the same seeded inputs.  Integer results (keys, hash array, vote
codes/counts, argmax) must be bit-exact; poses are compared bit-exact where
the device evaluates the same float expressions as the oracle, and within
1e-5 (absolute, per matrix entry) for the host greedy clustering average.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import ppf_amd
import ppf_oracle as O
import synth

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def medium():
    w = synth.plain_small(seed=3, n_model=500, n_clutter=700)
    return w.models[0], w.scenes[0], w.d_dists[0]


@pytest.mark.parametrize("op", range(9))
def test_device_numerics_match_host(ctx, op):
    rng = np.random.default_rng(op)
    n = 1 << 18
    if op == 0:
        a = rng.uniform(-1.05, 1.05, n).astype(np.float32)
        b = None
    elif op in (2, 3):
        a = rng.uniform(-3.3, 3.3, n).astype(np.float32)
        b = None
    elif op == 7:
        a = np.abs(rng.normal(size=n) * np.exp2(rng.integers(-30, 30, n))).astype(np.float32)
        b = None
    elif op == 8:
        a = rng.normal(size=2 * n).astype(np.float32)
        b = rng.normal(size=2 * n).astype(np.float32)
        a[:64] = np.array([1, 0, 1, -0.0, 0, 1, -0.0, 0] * 8, np.float32)
        b[:64] = np.array([-1, 0, -1, -0.0, 0, -1, 0, -1] * 8, np.float32)
    else:
        a = (rng.normal(size=n) * np.exp2(rng.integers(-12, 12, n))).astype(np.float32)
        b = (rng.normal(size=n) * np.exp2(rng.integers(-12, 4, n))).astype(np.float32)
        if op in (4, 5):
            a = np.abs(a) if op == 5 else a
            b = np.abs(b) + np.float32(1e-3)
    dev = ctx.debug_numerics(op, a, b)
    host = O.numerics(op, a, b)
    assert np.array_equal(_bits(dev), _bits(host)), f"op {op}: {np.sum(_bits(dev) != _bits(host))} differ"


def test_model_hash_array_matches_oracle(ctx, medium):
    m, _, d = medium
    model = ppf_amd.Model(m, d, ctx=ctx)
    hk, cnt, first, k2d = model.hash_array()
    keys = O.cloud_keys(m[0], m[1], 1, d)
    ok, oc, of, ok2d = O.hash_array(keys)
    assert np.array_equal(hk, ok)
    assert np.array_equal(cnt, oc.astype(np.uint32))
    assert np.array_equal(first, of.astype(np.uint32))
    assert np.array_equal(k2d, ok2d.astype(np.uint32))
    info = model.info()
    assert info["n_unique_keys"] == len(ok) and info["n_pairs"] == len(m[0]) ** 2


def test_model_hash_array_ranked_sort_matches_oracle(ctx):
    """A 1,500-point model (2.25e6 pairs): the build sorts its pair words by
    the ranks of the distinct keys; the hash array and pair map are the
    oracle's bit for bit."""
    w = synth.plain_small(seed=5, n_model=1500, n_clutter=10)
    m, d = w.models[0], w.d_dists[0]
    model = ppf_amd.Model(m, d, ctx=ctx)
    hk, cnt, first, k2d = model.hash_array()
    keys = O.cloud_keys(m[0], m[1], 1, d)
    ok, oc, of, ok2d = O.hash_array(keys)
    assert np.array_equal(hk, ok)
    assert np.array_equal(cnt, oc.astype(np.uint32))
    assert np.array_equal(first, of.astype(np.uint32))
    assert np.array_equal(k2d, ok2d.astype(np.uint32))


@pytest.mark.parametrize("case", ["one", "few_with_extremes", "model_like", "at_limit", "over_limit", "short"])
def test_device_key_sort_ranked(ctx, case):
    """The model build's sort of pair words by their high 32 bits: by ranks
    of the distinct keys up to 8,192 of them, by the radix sort beyond (and
    below 2^20 words); stable either way (random low halves)."""
    cases = ["one", "few_with_extremes", "model_like", "at_limit", "over_limit", "short"]
    rng = np.random.default_rng(100 + cases.index(case))
    n, U = {"one": (1 << 20, 1), "few_with_extremes": ((1 << 20) + 12345, 300),
            "model_like": ((1 << 22) + 777, 4200), "at_limit": ((1 << 21) + 3, 8192),
            "over_limit": ((1 << 21) + 3, 8193), "short": ((1 << 20) - 1, 100)}[case]
    distinct = np.unique(rng.integers(0, 2**32, 4 * U, dtype=np.uint64))
    distinct = rng.permutation(distinct)[:U]
    if case == "few_with_extremes":
        distinct[:2] = [0, 0xFFFFFFFF]
    hi = distinct[rng.integers(0, U, n)]
    hi[:U] = distinct  # every distinct key present
    hi = rng.permutation(hi)
    words = (hi << np.uint64(32)) | rng.integers(0, 2**32, n, dtype=np.uint64)
    out, ranked = ctx.debug_key_sort(words)
    assert ranked == (case not in ("over_limit", "short"))
    order = np.argsort(words >> np.uint64(32), kind="stable")
    assert np.array_equal(out, words[order])


def test_large_model_hash_array_full_size(ctx):
    """The 30,000-point model (9.0e8 pairs, the large-hash workload) at full
    size: the hash array's invariants (every pair counted once, keys strictly
    increasing, first = exclusive scan of counts) and, for sampled pairs, the
    oracle's key of the pair is one of the device's keys."""
    w = synth.large_hash()
    pts, nrm = w.models[0]
    d = w.d_dists[0]
    n = len(pts)
    model = ppf_amd.Model(w.models[0], d, ctx=ctx)
    hk, cnt, first, _ = model.hash_array(with_pair_map=False)
    assert model.info()["n_pairs"] == n * n
    assert np.all(hk[1:] > hk[:-1])
    assert np.all(cnt >= 1)
    assert int(cnt.astype(np.int64).sum()) == n * n
    excl = np.concatenate([[0], np.cumsum(cnt.astype(np.int64))[:-1]])
    assert np.array_equal(first.astype(np.int64), excl)
    if hk[0] == 0:
        assert cnt[0] >= n  # the diagonal has key 0
    rng = np.random.default_rng(11)
    for _ in range(300):
        i, j = (int(x) for x in rng.integers(0, n, 2))
        if i == j:
            continue
        _, q = O.compute_ppf(pts[i], nrm[i], pts[j], nrm[j], d)
        key = O.feature_key(q)
        pos = int(np.searchsorted(hk, key))
        assert pos < len(hk) and hk[pos] == key, (i, j, key)


def test_scene_keys_match_oracle(ctx, medium):
    _, s, d = medium
    scene = ppf_amd.Scene(s, d, 5, ctx=ctx)
    assert np.array_equal(scene.getHashKeys(d), O.cloud_keys(s[0], s[1], 5, d))


@pytest.mark.parametrize("exact", [False, True])
def test_votes_match_oracle(ctx, medium, exact):
    m, s, d = medium
    model = ppf_amd.Model(m, d, 0.3, ctx=ctx)
    scene = ppf_amd.Scene(s, d, 5, ctx=ctx)
    res = model.ppf_lookup(scene, exact_votes=exact)
    codes, counts = res.votes()
    oc, on, st = O.vote(m, s, d, 5, 0.3)
    assert res.info["max_count"] == st["max_count"]
    assert res.info["lookups"] == st["lookups"]
    assert np.array_equal(codes, oc) and np.array_equal(counts, on)


@pytest.fixture(scope="module")
def multi_tile():
    w = synth.plain_small(seed=5, n_model=1300, n_clutter=900)
    return w.models[0], w.scenes[0], w.d_dists[0]


def test_multi_tile_votes_match_oracle(ctx, multi_tile):
    # 1,300 model points in 3 tiles: the vote order placed from run lengths
    m, s, d = multi_tile
    model = ppf_amd.Model(m, d, 0.1, ctx=ctx)
    assert model.info()["n_tiles"] == 3
    res = model.ppf_lookup(ppf_amd.Scene(s, d, 5, ctx=ctx))
    codes, counts = res.votes()
    oc, on, st = O.vote(m, s, d, 5, 0.1)
    assert res.info["votes"] == st["votes"] and res.info["max_count"] == st["max_count"]
    assert np.array_equal(codes, oc) and np.array_equal(counts, on)


@pytest.mark.parametrize("exact", [False, True])
def test_per_reference_argmax_matches_oracle(ctx, multi_tile, exact):
    """Score pass (include/kernel.h:78-82): one bin per reference row, the
    maximum over all model-point tiles (count desc, code asc), then the usual
    threshold, poses and clustering."""
    m, s, d = multi_tile
    model = ppf_amd.Model(m, d, 0.3, ctx=ctx)
    assert model.info()["n_tiles"] > 1
    scene = ppf_amd.Scene(s, d, 5, ctx=ctx)
    res = model.ppf_lookup(scene, exact_votes=exact, per_reference_argmax=True)
    codes, counts = res.votes()
    oc, on, st = O.vote(m, s, d, 5, 0.3, per_ref=True)
    assert len(oc) > 10
    assert res.info["max_count"] == st["max_count"]
    assert np.array_equal(codes, oc) and np.array_equal(counts, on)
    assert len(np.unique(codes >> np.uint64(32))) == len(codes)
    opose, _ = O.register(m, s, d, 5, 0.3, False, per_ref=True)
    assert np.array_equal(_bits(res.pose()), _bits(opose))
    res.close()
    model.close()
    scene.close()


def test_tiny_golden_case(ctx):
    g = np.load(os.path.join(GOLDEN, "tiny_case.npz"))
    m = (g["model_points"], g["model_normals"])
    s = (g["scene_points"], g["scene_normals"])
    d = float(g["d_dist"])
    model = ppf_amd.Model(m, d, float(g["thr"]), ctx=ctx)
    res = model.ppf_lookup(ppf_amd.Scene(s, d, int(g["rpdf"]), ctx=ctx))
    codes, counts = res.votes()
    assert np.array_equal(codes, g["codes"]) and np.array_equal(counts, g["counts"])
    assert np.array_equal(_bits(res.transformations()), _bits(g["transforms"]))
    assert np.array_equal(_bits(res.pose()), _bits(g["pose"]))


def test_poses_and_spatial_clustering_match_oracle(ctx, medium):
    m, s, d = medium
    model = ppf_amd.Model(m, d, 0.3, ctx=ctx)
    res = model.ppf_lookup(ppf_amd.Scene(s, d, 5, ctx=ctx))
    codes, counts = res.votes()
    T = res.transformations()
    assert np.array_equal(_bits(T), _bits(O.poses(codes, m, s)))
    vc, tr, q = res.clusters()
    ref = O.cluster_spatial(T, codes, counts, d)
    assert np.array_equal(_bits(q), _bits(ref["quats"]))
    assert np.array_equal(_bits(vc), _bits(ref["vote_counts"]))
    assert np.array_equal(_bits(tr), _bits(ref["trans"]))
    assert res.info["max_idx"] == ref["max_idx"]
    opose, st = O.register(m, s, d, 5, 0.3, False)
    assert np.array_equal(_bits(res.pose()), _bits(opose))


@pytest.mark.parametrize("use_l1,use_avg", [(True, False), (False, True), (True, True)])
def test_clustering_flags_match_oracle(ctx, medium, use_l1, use_avg):
    # use_l1_norm drops the translation test; use_averaged_clusters averages
    # the translations of the passing members (kernel.cu:735-756)
    m, s, d = medium
    model = ppf_amd.Model(m, d, 0.3, use_l1_norm=use_l1, use_averaged_clusters=use_avg, ctx=ctx)
    res = model.ppf_lookup(ppf_amd.Scene(s, d, 5, ctx=ctx))
    codes, counts = res.votes()
    T = res.transformations()
    vc, tr, q = res.clusters()
    ref = O.cluster_spatial(T, codes, counts, d, use_l1=use_l1, use_avg=use_avg)
    assert np.array_equal(_bits(vc), _bits(ref["vote_counts"]))
    assert np.array_equal(_bits(tr), _bits(ref["trans"]))
    assert res.info["max_idx"] == ref["max_idx"]
    opose, _ = O.register(m, s, d, 5, 0.3, False, use_l1=use_l1, use_avg=use_avg)
    assert np.array_equal(_bits(res.pose()), _bits(opose))


def test_cpu_clustering_pose_matches_oracle(ctx, medium):
    m, s, d = medium
    model = ppf_amd.Model(m, d, 0.3, cpu_clustering=True, ctx=ctx)
    res = model.ppf_lookup(ppf_amd.Scene(s, d, 5, ctx=ctx))
    opose, st = O.register(m, s, d, 5, 0.3, True)
    assert res.info["n_clusters"] == st["cluster"]
    assert np.allclose(res.pose(), opose, atol=1e-5, rtol=0)
    # all of Model::cpu_transformations (up to three clusters) against the
    # oracle's greedy clustering of the same poses and counts
    _, counts = res.votes()
    d_angle0 = np.float32(np.float32(2.0 * np.float32(3.141592654)) / np.float32(30.0))
    oposes, ovotes, onc = O.cluster_greedy(res.transformations(), counts, d, d_angle0)
    poses, votes = res.cpu_transformations()
    assert onc == res.info["n_clusters"] and len(poses) == min(3, onc)
    assert votes.tolist() == ovotes.tolist()
    assert np.allclose(poses, oposes, atol=1e-5, rtol=0)


def test_plain_small_full_against_golden(ctx):
    with open(os.path.join(GOLDEN, "plain_small_oracle.json")) as f:
        gold = json.load(f)
    w = synth.plain_small()
    m, s, d = w.models[0], w.scenes[0], w.d_dists[0]
    model = ppf_amd.Model(m, d, 0.3, cpu_clustering=True, ctx=ctx)
    scene = ppf_amd.Scene(s, d, 5, ctx=ctx)
    res = model.ppf_lookup(scene)
    codes, counts = res.votes()
    h = hashlib.sha256(codes.tobytes() + counts.tobytes()).hexdigest()
    assert res.info["max_count"] == gold["stats"]["max_count"]
    assert res.info["n_kept"] == gold["stats"]["kept"]
    assert h == gold["votes_sha256"]
    assert hashlib.sha256(res.transformations().tobytes()).hexdigest() == gold["poses_sha256"]
    assert res.info["n_clusters"] == gold["cpu_n_clusters"]
    assert np.allclose(res.pose(), np.array(gold["pose_cpu_clustering"]), atol=1e-5, rtol=0)
    g2 = ppf_amd.Model(m, d, 0.3, cpu_clustering=False, ctx=ctx).ppf_lookup(scene)
    assert g2.info["max_idx"] == gold["gpu_max_idx"]
    assert np.array_equal(g2.pose(), np.array(gold["pose_gpu_clustering"], np.float32))


def test_fast_bins_equal_exact_bins_dataset_like(ctx):
    # full-size scene of the dataset-like workload, one model: the fast alpha
    # path must give the same kept votes as evaluating every vote exactly
    w = synth.dataset_like()
    m, s, d = w.models[0], w.scenes[0], w.d_dists[0]
    model = ppf_amd.Model(m, d, 0.3, ctx=ctx)
    scene = ppf_amd.Scene(s, d, 25, ctx=ctx)
    fast = model.ppf_lookup(scene)
    cf, nf = fast.votes()
    exact = model.ppf_lookup(scene, exact_votes=True)
    ce, ne = exact.votes()
    assert fast.info["max_count"] == exact.info["max_count"]
    assert np.array_equal(cf, ce) and np.array_equal(nf, ne)
    # run twice: identical
    again = model.ppf_lookup(scene)
    c2, n2 = again.votes()
    assert np.array_equal(cf, c2) and np.array_equal(nf, n2)
    # every kept pose maps its model reference point onto its scene point
    T = fast.transformations()
    for code, M in zip(cf[:200], T[:200]):
        sr, mr = int(code) >> 32, (int(code) & 0xFFFFFFFF) >> 6
        assert np.allclose(M[:3, :3] @ m[0][mr] + M[:3, 3], s[0][sr], atol=1e-3)


def test_fast_bins_equal_exact_bins_every_bin(ctx):
    # threshold 0 keeps every (s_r, m_r, alpha) bin with a vote: a vote put in
    # a neighbouring bin by the fast path would change two counts here
    w = synth.dataset_like()
    m, s, d = w.models[0], w.scenes[0], w.d_dists[0]
    model = ppf_amd.Model(m, d, 0.0, ctx=ctx)
    scene = ppf_amd.Scene(s, d, 400, ctx=ctx)
    fast = model.ppf_lookup(scene)
    exact = model.ppf_lookup(scene, exact_votes=True)
    cf, nf = fast.votes()
    ce, ne = exact.votes()
    assert fast.info["votes"] == exact.info["votes"] > 10**9
    assert len(cf) > 10**6
    assert np.array_equal(cf, ce) and np.array_equal(nf, ne)


def test_registration_api_matches_lookups(ctx):
    w1 = synth.plain_small(seed=5, n_model=300, n_clutter=300)
    w2 = synth.plain_small(seed=6, n_model=250, n_clutter=400)
    models = [w1.models[0], w2.models[0]]
    scenes = [w1.scenes[0], w2.scenes[0]]
    dd = [w1.d_dists[0], w2.d_dists[0]]
    poses, status = ppf_amd.ppf_registration(scenes, models, dd, 3, 0.3, False, ctx=ctx,
                                             return_status=True)
    for i, s in enumerate(scenes):
        for j, m in enumerate(models):
            res = ppf_amd.Model(m, dd[j], 0.3, ctx=ctx).ppf_lookup(ppf_amd.Scene(s, dd[j], 3, ctx=ctx))
            assert np.array_equal(poses[i][j], res.pose())
            assert status[i, j] == res.info["status"]


def test_model_weights_scale_cluster_votes(ctx, medium):
    m, s, d = medium
    rng = np.random.default_rng(4)
    wts = rng.uniform(0.5, 2.0, len(m[0])).astype(np.float32)
    model = ppf_amd.Model(m, d, 0.3, ctx=ctx)
    model.SetModelPointVoteWeights(wts)
    res = model.ppf_lookup(ppf_amd.Scene(s, d, 5, ctx=ctx))
    codes, counts = res.votes()
    vc, _, _ = res.clusters()
    ref = O.cluster_spatial(res.transformations(), codes, counts, d, weights=wts)
    assert np.array_equal(_bits(vc), _bits(ref["vote_counts"]))
    assert res.info["max_idx"] == ref["max_idx"]


def _line_model(seed=4):
    """Ellipsoid/box points plus a row of points on the x axis with normals
    (1, 0, 0): the pairs of that row have u = (dx, 0, 0) in the reference
    point's frame, so their records always take the float path."""
    m = synth.ellipsoid_box_model(80, seed)
    lp = np.zeros((14, 3), np.float32)
    lp[:, 0] = np.linspace(-0.3, 0.35, 14, dtype=np.float32)
    ln = np.tile(np.array([1, 0, 0], np.float32), (14, 1))
    return synth.as_cloud(np.concatenate([m[0], lp]), np.concatenate([m[1], ln]))


@pytest.mark.parametrize("exact", [False, True])
def test_votes_with_float_path_records_match_oracle(ctx, exact):
    m = _line_model()
    rng = np.random.default_rng(8)
    R = synth.random_rotation(rng)
    sp, sn = synth.transform(m[0].astype(np.float64), m[1].astype(np.float64), R, np.array([0.2, -0.1, 0.3]))
    cp = rng.uniform(-0.6, 0.6, (60, 3))
    cn = rng.normal(size=(60, 3))
    cn /= np.linalg.norm(cn, axis=1, keepdims=True)
    s = synth.as_cloud(np.concatenate([sp, cp]), np.concatenate([sn, cn]))
    d = float(np.float32(0.05))
    model = ppf_amd.Model(m, d, 0.0, ctx=ctx)
    res = model.ppf_lookup(ppf_amd.Scene(s, d, 1, ctx=ctx), exact_votes=exact)
    codes, counts = res.votes()
    oc, on, st = O.vote(m, s, d, 1, 0.0)
    assert st["votes"] > 0
    assert np.array_equal(codes, oc) and np.array_equal(counts, on)
    # the line's reference points have votes (their row pairs are float-path records)
    line_rows = set(range(80, 94))
    m_r = (codes & np.uint64(0xffffffff)) >> np.uint64(6)
    assert len(line_rows & set(m_r.tolist())) > 0


def test_degenerate_inputs(ctx):
    m = synth.ellipsoid_box_model(50, 1)
    far = (m[0] * 0 + np.float32(1000.0), m[1])
    model = ppf_amd.Model(m, 0.05, 0.3, ctx=ctx)
    # two coincident points far away: every key misses -> no votes, defined status
    s = (np.full((2, 3), 1000.0, np.float32), np.tile(np.array([0, 0, 1], np.float32), (2, 1)))
    res = model.ppf_lookup(ppf_amd.Scene(s, 0.05, 1, ctx=ctx))
    assert res.info["status"] == ppf_amd.PPF_ERR_NO_VOTES and res.n_kept == 0
    with pytest.raises(ppf_amd.PPFError):
        ppf_amd.Model((m[0][:1], m[1][:1]), 0.05, ctx=ctx)
    with pytest.raises(ppf_amd.PPFError):
        ppf_amd.Model(m, -1.0, ctx=ctx)
    del far


def test_budget_rejects_allocation():
    c = ppf_amd.Context(0, hbm_budget_bytes=4096)
    m = synth.ellipsoid_box_model(200, 2)
    with pytest.raises(ppf_amd.PPFError) as e:
        ppf_amd.Model(m, 0.05, ctx=c)
    assert e.value.code == ppf_amd.PPF_ERR_NOMEM
    c.close()


@pytest.mark.parametrize("n", [1, 1000, 4096 * 4096 + 12345])
def test_device_exclusive_scan(ctx, n):
    rng = np.random.default_rng(n)
    a = rng.integers(0, 4, n).astype(np.uint32)
    out, tot = ctx.debug_scan(a)
    want = np.concatenate([[0], np.cumsum(a, dtype=np.uint64)[:-1]]).astype(np.uint32)
    assert np.array_equal(out, want) and tot == int(a.sum())


@pytest.mark.parametrize("case", ["random", "equal", "few", "descending"])
def test_device_radix_sort_is_stable(ctx, case):
    rng = np.random.default_rng(1)
    n = 300001
    if case == "random":
        k = rng.integers(0, 2**63, n, dtype=np.uint64)
    elif case == "equal":
        k = np.full(n, 7, np.uint64)
    elif case == "few":
        k = rng.integers(0, 5, n).astype(np.uint64) << np.uint64(32)
    else:
        k = np.arange(n, dtype=np.uint64)[::-1].copy() << np.uint64(20)
    v = np.arange(n, dtype=np.uint32)
    for lo, hi in ((0, 64), (32, 64), (20, 44)):
        sk, sv = ctx.debug_sort(k, v, lo, hi)
        digit = (k >> np.uint64(lo)) & np.uint64((1 << (hi - lo)) - 1) if hi - lo < 64 else k
        order = np.argsort(digit, kind="stable")
        assert np.array_equal(sv, v[order]) and np.array_equal(sk, k[order])


@pytest.mark.parametrize("shape", [(256, 16), (256, 8), (512, 8), (512, 16)])
@pytest.mark.parametrize("n", [2, 2047, 2048, 2049, 4095, 4096, 4097, 8193, 1 << 21])
def test_device_radix_sort_sizes(n, shape):
    """Tile edges of the one-sweep sort (threads x keys per thread per tile),
    keys alone and with values."""
    c = ppf_amd.Context(0)
    try:
        c.debug_set(2, shape[0])
        c.debug_set(1, shape[1])
        rng = np.random.default_rng(n)
        k = rng.integers(0, 2**40, n, dtype=np.uint64)
        k[::3] = k[0]  # many equal keys
        sk, _ = c.debug_sort(k, None, 0, 40)
        assert np.array_equal(sk, np.sort(k, kind="stable"))
        v = np.arange(n, dtype=np.uint32)
        sk, sv = c.debug_sort(k, v, 16, 32)
        order = np.argsort((k >> np.uint64(16)) & np.uint64(0xffff), kind="stable")
        assert np.array_equal(sk, k[order]) and np.array_equal(sv, v[order])
    finally:
        c.close()


def test_device_radix_sort_big_tiles(ctx):
    """Sorts of 2^24 keys and more use 512 x 16-key tiles; a partial last tile."""
    n = (1 << 24) + 12345
    rng = np.random.default_rng(24)
    k = rng.integers(0, 2**48, n, dtype=np.uint64)
    k[: n // 2] &= np.uint64(0xFFFF0000FFFF)  # runs of equal digits in the middle bytes
    v = np.arange(n, dtype=np.uint32)
    sk, sv = ctx.debug_sort(k, v, 8, 40)
    order = np.argsort((k >> np.uint64(8)) & np.uint64(0xFFFFFFFF), kind="stable")
    assert np.array_equal(sv, v[order]) and np.array_equal(sk, k[order])


def test_device_copy_kernel(ctx):
    a = np.random.default_rng(0).integers(0, 2**32, 1 << 20, dtype=np.uint64).astype(np.uint32)
    assert np.array_equal(ctx.debug_copy(a), a)


def test_candidate_overflow_reruns_give_same_result(ctx):
    """A candidate buffer far too small for the kept votes forces a rerun of
    the vote pass, sized to the first pass's candidate count (rounded up);
    results equal the default-capacity lookup, in the plain, batched and
    served forms."""
    w = synth.plain_small(seed=13, n_model=500, n_clutter=800)
    m = ppf_amd.Model(w.models[0], w.d_dists[0], w.vote_count_threshold, ctx=ctx)
    s = ppf_amd.Scene(w.scenes[0], 0.0, w.rpdf, ctx=ctx)
    ref = m.ppf_lookup(s)
    assert ref.info["n_kept"] > 600  # > 512: the 257-entry buffer (rounded to 512) overflows
    c0, n0 = ref.votes()
    small = m.ppf_lookup(s, candidate_capacity=257)
    c1, n1 = small.votes()
    assert np.array_equal(c0, c1) and np.array_equal(n0, n1)
    assert small.pose().tobytes() == ref.pose().tobytes()
    assert small.info["votes"] == ref.info["votes"]
    res, info = ppf_amd.batch_lookup([m, m], [s, s], w.vote_count_threshold, ctx=ctx,
                                     candidate_capacity=300)
    assert info["vote_launches"] >= 2
    for r in res:
        c2, n2 = r.votes()
        assert np.array_equal(c0, c2) and np.array_equal(n0, n2)
        assert r.pose().tobytes() == ref.pose().tobytes()
        r.close()
    srv = ppf_amd.Server([m], round_ms=1.0, ctx=ctx)
    try:
        t = srv.submit(w.scenes[0], [0], w.rpdf, w.vote_count_threshold, candidate_capacity=100)
        poses, st, info, out = srv.wait(t, 120.0, results=True)
        assert info["rc"] == ppf_amd.PPF_OK
        c3, n3 = out[0].votes()
        assert np.array_equal(c0, c3) and np.array_equal(n0, n3)
        assert poses[0].tobytes() == ref.pose().tobytes()
        out[0].close()
    finally:
        srv.close()
    for o in (small, ref, m, s):
        o.close()


def test_progress_reports_and_sliced_votes_give_same_result():
    """ppf_ctx_set_progress: voting in launches of 37 reference rows (running
    maximum and candidates carried over) gives the one-launch result bit for
    bit, with a report per launch and per later stage."""
    w = synth.plain_small(seed=15, n_model=500, n_clutter=900)
    c = ppf_amd.Context(0)
    try:
        m = ppf_amd.Model(w.models[0], w.d_dists[0], 0.3, ctx=c)
        s = ppf_amd.Scene(w.scenes[0], 0.0, w.rpdf, ctx=c)
        ref = m.ppf_lookup(s)
        rc, rn = ref.votes()
        seen = []
        c.set_progress(lambda stage, done, total: seen.append((stage, done, total)), rows_per_launch=37)
        res = m.ppf_lookup(s)
        c.set_progress(None)
        codes, counts = res.votes()
        assert np.array_equal(codes, rc) and np.array_equal(counts, rn)
        assert res.pose().tobytes() == ref.pose().tobytes()
        R = (len(w.scenes[0][0]) + w.rpdf - 1) // w.rpdf
        votes = [d for st, d, t in seen if st == "vote"]
        assert votes == list(range(37, R, 37)) + [R]
        assert [st for st, _, _ in seen if st != "vote"] == ["select", "poses", "cluster"]
    finally:
        c.close()
