"""CPU pins of the oracle's pipeline stages (oracle/ppf_oracle.cpp).

Each stage is checked against something computed a different way: numpy's
stable argsort/unique for the hash array, a Python dictionary join for the vote
histogram, geometric invariants for poses, an O(P^2) scan for the spatial-hash
clustering, and the generator's ground truth for whole registrations.
"""
import collections
import math
import struct

import numpy as np
import pytest

import ppf_oracle as O
import synth


def tiny_workload(seed=0, n_model=40, n_scene=70, noise=True):
    rng = np.random.default_rng(seed)
    m = synth.ellipsoid_box_model(n_model, seed + 5)
    R = synth.random_rotation(rng)
    t = rng.uniform(-1, 1, 3)
    sp, sn = synth.transform(m[0].astype(np.float64), m[1].astype(np.float64), R, t)
    if noise:
        sp, sn = synth.perturb(sp, sn, 0.002, 2.0, rng)
    k = n_scene - n_model
    cp = rng.uniform(-1.5, 1.5, (k, 3)) + t
    cn = synth._normalize(rng.normal(size=(k, 3)))
    P = np.concatenate([sp, cp])
    N = np.concatenate([sn, cn])
    perm = rng.permutation(len(P))
    D = synth.bbox_diameter(m[0])
    return m, synth.as_cloud(P[perm], N[perm]), float(np.float32(0.1 * D)), synth.pose_matrix(R, t)


def test_hash_array_matches_numpy():
    rng = np.random.default_rng(0)
    for keys in (rng.integers(0, 50, 3000).astype(np.uint32),
                 np.zeros(100, np.uint32),
                 np.arange(500, dtype=np.uint32)[::-1].copy(),
                 np.array([7], np.uint32)):
        hk, cnt, first, k2d = O.hash_array(keys)
        order = np.argsort(keys, kind="stable")
        u, c = np.unique(keys, return_counts=True)
        assert np.array_equal(k2d, order)
        assert np.array_equal(hk, u) and np.array_equal(cnt, c)
        assert np.array_equal(first, np.concatenate([[0], np.cumsum(c)[:-1]]))


def test_cloud_keys_rows_and_diagonal():
    m, s, d, _ = tiny_workload()
    keys = O.cloud_keys(s[0], s[1], 3, d).reshape(len(s[0]), -1)
    n = len(s[0])
    for r in range(n):
        if r % 3:
            assert not keys[r].any()
        else:
            assert keys[r, r] == 0
    # a key equals the FNV-1a of the quantized feature bytes
    raw, q = O.compute_ppf(s[0][3], s[1][3], s[0][5], s[1][5], d)
    assert keys[3, 5] == O.fnv1a(q.tobytes())


def _python_votes(model, scene, d, rpdf, thr):
    """Dictionary join of scene and model keys, alpha from the literal
    trans_model_scene, Counter histogram, then the reference threshold/order."""
    mp, mn = model
    sp, sn = scene
    n_m, n_s = len(mp), len(sp)
    mk = O.cloud_keys(mp, mn, 1, d).reshape(n_m, n_m)
    sk = O.cloud_keys(sp, sn, rpdf, d).reshape(n_s, n_s)
    table = collections.defaultdict(list)
    for r in range(n_m):
        for i in range(n_m):
            table[int(mk[r, i])].append((r, i))
    hist = collections.Counter()
    for r in range(0, n_s, rpdf):
        for i in range(n_s):
            k = int(sk[r, i])
            if k == 0 or k not in table:
                continue
            for (mr, mi) in table[k]:
                a = O.alpha_literal(mp[mr], mn[mr], mp[mi], sp[r], sn[r], sp[i])
                hist[(r << 32) | (mr << 6) | a] += 1
    items = sorted(hist.items())                                   # code ascending
    items.sort(key=lambda kv: -np.float32(kv[1]))                  # stable, count desc
    if not items:
        return np.zeros(0, np.uint64), np.zeros(0, np.uint32)
    thr_f = np.float32(np.float32(thr) * np.float32(items[0][1]))
    kept = [(c, n) for c, n in items if np.float32(n) > thr_f]
    return (np.array([c for c, _ in kept], np.uint64), np.array([n for _, n in kept], np.uint32))


def test_votes_match_python_join():
    m, s, d, _ = tiny_workload(seed=1, n_model=30, n_scene=50)
    codes, counts, st = O.vote(m, s, d, 3, 0.3, mode=0)
    pc, pn = _python_votes(m, s, d, 3, 0.3)
    assert st["votes"] > 0
    assert np.array_equal(codes, pc) and np.array_equal(counts, pn)


@pytest.mark.parametrize("seed", [2, 3])
def test_literal_and_hoisted_votes_identical(seed):
    m, s, d, _ = tiny_workload(seed=seed, n_model=60, n_scene=120)
    c0, n0, s0 = O.vote(m, s, d, 2, 0.2, mode=0)
    c1, n1, s1 = O.vote(m, s, d, 2, 0.2, mode=1)
    assert s0 == s1
    assert np.array_equal(c0, c1) and np.array_equal(n0, n1)


def test_vote_ref_subset_is_a_restriction():
    m, s, d, _ = tiny_workload(seed=4, n_model=50, n_scene=90)
    c_all, n_all, st_all = O.vote(m, s, d, 1, 0.0)
    refs = np.array([4, 9, 30], np.int64)
    c_sub, n_sub, st_sub = O.vote(m, s, d, 1, 0.0, refs=refs)
    sub = set(zip(c_sub.tolist(), n_sub.tolist()))
    want = {(c, n) for c, n in zip(c_all.tolist(), n_all.tolist()) if (c >> 32) in set(refs.tolist())}
    assert sub == want
    assert st_sub["lookups"] == len(refs) * (len(s[0]) - 1)


@pytest.mark.parametrize("seed", [6, 7])
def test_per_ref_argmax_is_grouped_numpy_argmax(seed):
    """The per-reference-point score pass against numpy: from the full vote
    list (threshold 0), each s_r's bin with the largest count, smallest code on
    ties (lexsort), then the count threshold against the global maximum and the
    (count desc, code asc) order of the thresholded list."""
    m, s, d, _ = tiny_workload(seed=seed, n_model=70, n_scene=160)
    c_all, n_all, st_all = O.vote(m, s, d, 2, 0.0)
    order = np.lexsort((c_all, -n_all.astype(np.int64), c_all >> np.uint64(32)))
    sr = (c_all >> np.uint64(32))[order]
    first = np.concatenate([[True], sr[1:] != sr[:-1]])
    bc, bn = c_all[order][first], n_all[order][first]
    assert len(np.unique(c_all >> np.uint64(32))) == len(bc)
    for thr in (0.0, 0.3, 0.7):
        keep = bn.astype(np.float32) > np.float32(thr) * np.float32(bn.max())
        kc, kn = bc[keep], bn[keep]
        o = np.lexsort((kc, -kn.astype(np.int64)))
        c, n, st = O.vote(m, s, d, 2, thr, per_ref=True)
        assert np.array_equal(c, kc[o]) and np.array_equal(n, kn[o])
        assert st["max_count"] == st_all["max_count"] == int(bn.max())
    # the per-reference list is what the full list keeps at one bin per row
    assert set(zip(bc.tolist(), bn.tolist())) <= set(zip(c_all.tolist(), n_all.tolist()))


def test_pose_invariants():
    m, s, d, _ = tiny_workload(seed=5, n_model=60, n_scene=100)
    codes, counts, _ = O.vote(m, s, d, 2, 0.0)
    T = O.poses(codes, m, s)
    for code, M in zip(codes[:500], T[:500]):
        if code == 0:
            assert not M.any()
            continue
        sr, mr = int(code) >> 32, (int(code) & 0xFFFFFFFF) >> 6
        R = M[:3, :3].astype(np.float64)
        assert np.allclose(R @ R.T, np.eye(3), atol=2e-5)
        assert abs(np.linalg.det(R) - 1) < 2e-5
        # T maps the model reference point onto the scene reference point and
        # its normal onto the scene normal (compute_transforms, kernel.cu:372-401)
        assert np.allclose(R @ m[0][mr] + M[:3, 3], s[0][sr], atol=1e-4)
        nm = m[1][mr] / np.linalg.norm(m[1][mr])
        ns = s[1][sr] / np.linalg.norm(s[1][sr])
        assert np.allclose(R @ nm, ns, atol=1e-4)
        assert np.allclose(M[3], [0, 0, 0, 1])


def test_exact_copy_top_vote_recovers_pose():
    # scene = model under a rigid motion, no noise, no clutter: the best vote
    # pose is within one angle bin of the truth
    m, s, d, truth = tiny_workload(seed=6, n_model=120, n_scene=120, noise=False)
    codes, counts, _ = O.vote(m, s, d, 1, 0.5)
    T = O.poses(codes[:1], m, s)[0].astype(np.float64)
    dR = T[:3, :3].T @ truth[:3, :3]
    ang = math.acos(max(-1.0, min(1.0, (np.trace(dR) - 1) / 2)))
    assert ang < 2 * 2 * math.pi / 30
    assert np.linalg.norm(T[:3, 3] - truth[:3, 3]) < 0.5


def _fnv_bytes(b):
    h = 0x811C9DC5
    for x in b:
        v = x if x < 128 else (x | 0xFFFFFF00)
        h = ((h ^ v) * 16777619) % 2**32
    return h


def test_spatial_clustering_bruteforce():
    rng = np.random.default_rng(7)
    P = 300
    base = [synth.random_rotation(rng) for _ in range(4)]
    T = np.zeros((P, 4, 4), np.float32)
    for i in range(P):
        R = base[i % 4] @ synth.random_rotation(np.random.default_rng(i)) if i % 7 == 0 else base[i % 4]
        T[i, :3, :3] = R
        T[i, :3, 3] = rng.normal(0, 0.3, 3) + (i % 4)
        T[i, 3, 3] = 1
    counts = rng.integers(1, 50, P).astype(np.uint32)
    codes = (np.arange(P, dtype=np.uint64) << np.uint64(6)) | np.uint64(3)
    d = 0.2
    out = O.cluster_spatial(T, codes, counts, d)
    # brute force: cells, FNV of the int32 cell, 26 neighbour hashes, O(P^2) scan
    tr = T[:, :3, 3]
    q = out["quats"]
    cells = []
    for t in tr:
        c = [int(np.trunc(np.float32(np.float32(x) - np.fmod(np.float32(x), np.float32(d))) /
                          np.float32(d))) for x in t]
        cells.append(c)
    hashes = [_fnv_bytes(struct.pack("<3i", *c)) for c in cells]
    assert np.array_equal(out["hashes"], np.array(hashes, np.uint32))
    rot_sq = np.float32(2 * np.float32(np.float32(2 * np.float32(3.141592654)) / np.float32(30))) ** 2
    for i in range(P):
        vc = 1.0
        for a in (-1, 0, 1):
            for b in (-1, 0, 1):
                for k in (-1, 0, 1):
                    if a == b == k == 0:
                        continue
                    h = _fnv_bytes(struct.pack("<3i", cells[i][0] + a, cells[i][1] + b, cells[i][2] + k))
                    for j in range(P):
                        if hashes[j] != h:
                            continue
                        qd = abs(8 * (1 - float(np.dot(q[i].astype(np.float64), q[j]))))
                        if qd < rot_sq and np.linalg.norm(tr[i] - tr[j]) < d:
                            vc += float(counts[j])
        assert abs(out["vote_counts"][i] - vc) < 1e-3, i
    assert out["max_idx"] == int(np.argmax(out["vote_counts"]))


def test_quaternion_of_rotation():
    rng = np.random.default_rng(8)
    for _ in range(50):
        R = synth.random_rotation(rng)
        T = np.eye(4, dtype=np.float32)
        T[:3, :3] = R
        out = O.cluster_spatial(T[None], np.array([64], np.uint64), np.array([1], np.uint32), 1.0)
        qv = out["quats"][0].astype(np.float64)
        # scalar first, unit length for an exact rotation (kernel.cu:128-144)
        w = math.sqrt(max(0.0, 1 + np.trace(R))) / 2
        assert abs(abs(qv[0]) - w) < 1e-5 and abs(np.linalg.norm(qv) - 1) < 1e-5
        x, y, z = qv[1:]
        Rq = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * qv[0]), 2 * (x * z + y * qv[0])],
                       [2 * (x * y + z * qv[0]), 1 - 2 * (x * x + z * z), 2 * (y * z - x * qv[0])],
                       [2 * (x * z - y * qv[0]), 2 * (y * z + x * qv[0]), 1 - 2 * (x * x + y * y)]])
        assert np.allclose(Rq, R, atol=1e-4)


def test_greedy_clustering_invariants():
    rng = np.random.default_rng(9)
    R = synth.random_rotation(rng)
    T = np.tile(np.eye(4, dtype=np.float32), (6, 1, 1))
    for i in range(6):
        T[i, :3, :3] = R
        T[i, :3, 3] = [0.0, 0.0, 0.0] if i < 4 else [5.0, 0.0, 0.0]
    counts = np.array([5, 4, 3, 2, 20, 1], np.uint32)
    poses, votes, nc = O.cluster_greedy(T, counts, 0.5, 0.2)
    assert nc == 2
    assert votes.tolist() == [21, 14]          # cluster sums, sorted descending
    assert np.allclose(poses[0][:3, 3], [5, 0, 0]) and np.allclose(poses[0][:3, :3], R, atol=1e-5)
    assert np.allclose(poses[1][:3, 3], [0, 0, 0]) and np.allclose(poses[1][:3, :3], R, atol=1e-5)


@pytest.mark.parametrize("cpu", [True, False])
def test_registration_recovers_ground_truth(cpu):
    w = synth.plain_small(seed=11, n_model=400, n_clutter=300)
    pose, st = O.register(w.models[0], w.scenes[0], w.d_dists[0], 2, 0.3, cpu)
    truth = w.truths[0][0]
    D = w.d_dists[0] / 0.05
    dR = pose[:3, :3].astype(np.float64).T @ truth[:3, :3]
    ang = math.degrees(math.acos(max(-1.0, min(1.0, (np.trace(dR) - 1) / 2))))
    # the reference's validation rule: translation < 0.1*diameter, rotation < 12 deg
    assert np.linalg.norm(pose[:3, 3] - truth[:3, 3]) < 0.1 * D
    assert ang < 12.0


def test_timing_form_counts_match_vote_histogram():
    """The row-parallel timing loop (used for the CPU baseline) does the same
    work as the vote histogram: same lookups, hits, votes and nonzero bins."""
    m, s, d, _ = tiny_workload(seed=5, n_model=80, n_scene=400)
    refs = np.array([0, 7, 133, 399], np.int64)
    a = O.vote_throughput(m, s, d, refs)
    om = O.OracleModel(m, d)
    b = om.vote_throughput(s, refs)
    for k in ("lookups", "hits", "votes", "entries"):
        assert a[k] == b[k], k
    assert a["votes"] > 0 and a["lookups"] == len(refs) * (len(s[0]) - 1)


# ---- discriminating pins of the two clustering paths' geometry
D_ANGLE0 = float(np.float32(np.float32(2.0 * np.float32(3.141592654)) / np.float32(30.0)))


def _rx(t):
    c, s = math.cos(t), math.sin(t)
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]], np.float64)


def _rz(t):
    c, s = math.cos(t), math.sin(t)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]], np.float64)


def _quat64(R):
    """Unit quaternion (w, x, y, z) of a float64 rotation (closed form)."""
    w = math.sqrt(max(0.0, 1.0 + np.trace(R))) / 2
    x = math.copysign(math.sqrt(max(0.0, 1 + R[0, 0] - R[1, 1] - R[2, 2])) / 2, R[2, 1] - R[1, 2])
    y = math.copysign(math.sqrt(max(0.0, 1 - R[0, 0] + R[1, 1] - R[2, 2])) / 2, R[0, 2] - R[2, 0])
    z = math.copysign(math.sqrt(max(0.0, 1 - R[0, 0] - R[1, 1] + R[2, 2])) / 2, R[1, 0] - R[0, 1])
    return np.array([w, x, y, z])


def _rot64(q):
    w, x, y, z = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


@pytest.mark.parametrize("theta", [D_ANGLE0 - 1e-3, D_ANGLE0 + 1e-3, 0.5, 3.0])
def test_rot_angle_between_closed_form(theta):
    """The pose test's angle of A^-1 B (transformation_clustering.cpp:131-133)
    equals theta for B = A Rx(theta) and B = Rz(theta) A, for any A: a product
    without the inverse, or the angle of one operand, gives another value."""
    rng = np.random.default_rng(31)
    for _ in range(20):
        A = synth.random_rotation(rng)
        for B in (A @ _rx(theta), _rz(theta) @ A):
            got = O.rot_angle_between(A.astype(np.float32), B.astype(np.float32))
            assert abs(got - theta) < 2e-5, (got, theta)
        # the angle is of A^-1 B, not of A B or B alone
        assert abs(O.rot_angle_between(A.astype(np.float32), (A @ _rx(theta)).astype(np.float32)) -
                   O.rot_angle_between(np.eye(3, dtype=np.float32), (A @ A @ _rx(theta)).astype(np.float32))) > 1e-3 \
            or abs(np.trace(A) - 3) < 1e-2


def _pose(R, t):
    T = np.eye(4, dtype=np.float32)
    T[:3, :3] = R
    T[:3, 3] = t
    return T


def test_greedy_compares_with_first_member_only():
    """transformation_clustering.cpp:74-92: a pose joins the first cluster whose
    FIRST pose is within bounds.  Pose 3 is within bounds of pose 2 (a member of
    cluster 1) but not of pose 1, so it opens cluster 2 -- by translation, and
    separately by rotation."""
    rng = np.random.default_rng(32)
    R = synth.random_rotation(rng)
    T = np.stack([_pose(R, [0, 0, 0]), _pose(R, [0.4, 0, 0]), _pose(R, [0.8, 0, 0])])
    _, votes, nc = O.cluster_greedy(T, np.array([10, 9, 8], np.uint32), 0.5, 0.2)
    assert nc == 2 and votes.tolist() == [19, 8]
    T = np.stack([_pose(R, [0, 0, 0]), _pose(R @ _rx(0.15), [0, 0, 0]), _pose(R @ _rx(0.30), [0, 0, 0])])
    _, votes, nc = O.cluster_greedy(T, np.array([10, 9, 8], np.uint32), 1.0, 0.2)
    assert nc == 2 and votes.tolist() == [19, 8]
    # the same chain in the opposite vote order: pose 2 (middle) comes first
    # and takes both neighbours
    _, votes, nc = O.cluster_greedy(T, np.array([8, 10, 9], np.uint32), 1.0, 0.2)
    assert nc == 1 and votes.tolist() == [27]


def _shepperd64(R):
    """(w, x, y, z) of a rotation matrix by Shepperd's method as Eigen's
    Quaternion(Matrix3) applies it: the trace branch when the trace is
    positive, else the branch of the largest diagonal element.  Signs are
    those of the branch (no canonical hemisphere)."""
    t = np.trace(R)
    if t > 0:
        s = math.sqrt(t + 1.0)
        w = 0.5 * s
        s = 0.5 / s
        return np.array([w, (R[2, 1] - R[1, 2]) * s, (R[0, 2] - R[2, 0]) * s, (R[1, 0] - R[0, 1]) * s])
    i = 0
    if R[1, 1] > R[0, 0]:
        i = 1
    if R[2, 2] > R[i, i]:
        i = 2
    j, k = (i + 1) % 3, (i + 2) % 3
    c = np.zeros(3)
    s = math.sqrt(R[i, i] - R[j, j] - R[k, k] + 1.0)
    c[i] = 0.5 * s
    s = 0.5 / s
    w = (R[k, j] - R[j, k]) * s
    c[j] = (R[j, i] + R[i, j]) * s
    c[k] = (R[k, i] + R[i, k]) * s
    return np.array([w, c[0], c[1], c[2]])


def test_greedy_average_is_unweighted_quaternion_mean():
    """transformation_clustering.cpp:105-120: translation mean and the
    normalised mean of the quaternion coefficients (Eigen's matrix->quaternion,
    no sign alignment), unweighted by votes.  Near the identity (trace > 0) the
    rotations Rx(0), Rx(0.1), Rx(0.2) average to Rx(0.1) exactly (half-angle
    quaternions symmetric about 0.05); a vote-weighted mean (10:1:1) would give
    ~Rx(0.025).  For random rotations the result equals the float64 mean of
    Shepperd quaternions, signs as each branch gives them."""
    rng = np.random.default_rng(33)
    for k in range(12):
        R = _rz(rng.uniform(-0.5, 0.5)) if k < 4 else synth.random_rotation(rng)
        ts = rng.uniform(-1, 1, (3, 3))
        Rs = [(R @ _rx(a)).astype(np.float32) for a in (0.0, 0.1, 0.2)]
        T = np.stack([_pose(Ri, t) for Ri, t in zip(Rs, ts)])
        poses, votes, nc = O.cluster_greedy(T, np.array([10, 1, 1], np.uint32), 5.0, 0.5)
        assert nc == 1 and votes.tolist() == [12]
        assert np.allclose(poses[0][:3, 3], ts.astype(np.float32).mean(axis=0), atol=1e-6)
        want = _rot64(sum(_shepperd64(Ri.astype(np.float64)) for Ri in Rs))
        assert np.allclose(poses[0][:3, :3], want, atol=3e-6)
        if k < 4:
            assert np.allclose(poses[0][:3, :3], R @ _rx(0.1), atol=3e-6)
            weighted = _rot64(sum(w * _shepperd64(Ri.astype(np.float64)) for w, Ri in zip((10, 1, 1), Rs)))
            assert not np.allclose(poses[0][:3, :3], weighted, atol=1e-3)
    # two rotations about different axes
    for _ in range(10):
        R1 = synth.random_rotation(rng).astype(np.float32)
        R2 = (R1.astype(np.float64) @ _rz(0.3) @ _rx(-0.2)).astype(np.float32)
        T = np.stack([_pose(R1, [0, 0, 0]), _pose(R2, [0, 0, 0])])
        poses, _, nc = O.cluster_greedy(T, np.array([3, 7], np.uint32), 5.0, 1.0)
        assert nc == 1
        want = _rot64(_shepperd64(R1.astype(np.float64)) + _shepperd64(R2.astype(np.float64)))
        assert np.allclose(poses[0][:3, :3], want, atol=3e-6)


def _hrotmat2quat_f32(T):
    """kernel.cu:128-144 in numpy float32, one IEEE operation at a time."""
    f = np.float32
    T = T.astype(np.float32)
    t = f(f(T[0, 0] + T[1, 1]) + T[2, 2])
    q = np.empty(4, np.float32)
    q[0] = f(0.5) * np.sqrt(f(1) + t)
    q[1] = np.copysign(f(0.5) * np.sqrt(f(f(f(1) + T[0, 0]) - T[1, 1]) - T[2, 2]), f(T[2, 1] - T[1, 2]))
    q[2] = np.copysign(f(0.5) * np.sqrt(f(f(f(1) - T[0, 0]) + T[1, 1]) - T[2, 2]), f(T[0, 2] - T[2, 0]))
    q[3] = np.copysign(f(0.5) * np.sqrt(f(f(f(1) - T[0, 0]) - T[1, 1]) + T[2, 2]), f(T[1, 0] - T[0, 1]))
    dot = f(f(f(q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]) + q[3] * q[3])
    return q, np.sqrt(dot)


def test_hrotmat2quat_divides_by_sqrt_of_norm():
    """hrotmat2quat divides by sqrt(norm(q)) = |q|^(1/2) (kernel.cu:138).  The
    four square-root arguments sum to 4 for ANY matrix, so |q| = 1 up to
    rounding and the quirk only shows in the low bits: on general 3x3 blocks
    (entries uniform in [-1, 1], as far from a rotation as the radicands
    allow) the oracle equals the float32 restatement of kernel.cu:128-144 bit
    for bit, near-rotations included, and dividing by |q| instead differs on
    some of them (~1 in 900)."""
    rng = np.random.default_rng(34)
    differs = 0
    for i in range(12000):
        T = np.eye(4, dtype=np.float32)
        if i % 4 == 0:
            T[:3, :3] = synth.random_rotation(rng) + rng.normal(scale=1e-3, size=(3, 3))
        else:
            T[:3, :3] = rng.uniform(-1, 1, (3, 3))
        with np.errstate(invalid="ignore"):
            q, nrm = _hrotmat2quat_f32(T)
        if not np.all(np.isfinite(q)):
            continue
        want = q / np.sqrt(nrm)
        got = O.hrotmat2quat(T)
        assert got.tobytes() == want.astype(np.float32).tobytes()
        differs += (q / nrm).astype(np.float32).tobytes() != want.tobytes()
    assert differs > 0


def test_spatial_clustering_shared_hash_cells_bruteforce():
    """Two distinct cells with one FNV hash (kernel.cu:663-699): the reference
    finds neighbours by hash, so a pose also counts members of the other cell
    whenever that cell's hash is one of its 26 neighbour hashes.  The oracle
    equals an O(P^2) numpy scan with that rule."""
    import collide
    T, codes, counts, d, (a, b) = collide.colliding_poses(seed=3)
    out = O.cluster_spatial(T, codes, counts, d)
    P = len(T)
    tr = T[:, :3, 3].astype(np.float32)
    cells = np.trunc(tr - np.fmod(tr, np.float32(d))).astype(np.int64)
    hashes = collide.fnv_cells(cells)
    assert np.array_equal(out["hashes"], hashes)
    assert len(set(hashes.tolist())) < len({tuple(c) for c in cells.tolist()})  # a shared hash
    q = out["quats"].astype(np.float64)
    rot_sq = np.float32(2 * np.float32(np.float32(2 * np.float32(3.141592654)) / np.float32(30))) ** 2
    offs = np.array([(x, y, z) for x in (-1, 0, 1) for y in (-1, 0, 1) for z in (-1, 0, 1)
                     if (x, y, z) != (0, 0, 0)])
    order = np.lexsort((np.arange(P), hashes))  # members in (hash, index) order
    for i in range(P):
        nh = collide.fnv_cells(cells[i] + offs)
        vc = 1.0
        for h in nh:
            for j in order[hashes[order] == h]:
                qd = abs(8 * (1 - float(np.dot(q[i], q[j]))))
                if qd < rot_sq and np.linalg.norm(tr[i].astype(np.float64) - tr[j]) < d:
                    vc += float(counts[j])
        assert abs(out["vote_counts"][i] - vc) < 1e-3, i
    # poses of the two colliding cells see each other only through the hash
    ia = np.nonzero((cells == a).all(1))[0]
    ib = np.nonzero((cells == b).all(1))[0]
    assert len(ia) and len(ib)
