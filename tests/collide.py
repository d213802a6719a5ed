"""Pose sets whose translation cells collide under the spatial hash: two
distinct integer cells with the same FNV-1a hash of their int32 triple
(kernel.cu:20-30, 663-699), found by a birthday search, with poses in both
cells and in their neighbours.  Used to run the clustering's shared-hash path
on the CPU oracle and on the GPU."""
import numpy as np

import synth


def fnv_cells(c):
    """FNV-1a (bytes as signed char) of int32 triples c [n][3] -> uint32 [n]."""
    b = np.ascontiguousarray(c, np.int32).view(np.uint8).reshape(len(c), 12)
    h = np.full(len(c), 0x811C9DC5, np.uint64)
    for i in range(12):
        v = b[:, i].astype(np.uint64)
        v = np.where(v >= 128, v | np.uint64(0xFFFFFF00), v)
        h = ((h ^ v) * np.uint64(16777619)) & np.uint64(0xFFFFFFFF)
    return h.astype(np.uint32)


def colliding_cells():
    """Two distinct cells with positive coordinates and equal hashes."""
    rng = np.random.default_rng(12345)
    g = np.unique(rng.integers(2, 1 << 19, (1 << 21, 3)), axis=0)
    h = fnv_cells(g)
    order = np.argsort(h, kind="stable")
    hs = h[order]
    k = np.nonzero(hs[1:] == hs[:-1])[0][0]
    a, b = g[order[k]], g[order[k + 1]]
    assert fnv_cells(np.stack([a, b]))[0] == fnv_cells(np.stack([a, b]))[1] and not np.array_equal(a, b)
    return a, b


def colliding_poses(seed=0, per_cell=150, d_dist=1.0):
    """(T [P][4][4], codes, counts, d_dist, cells): poses in two colliding cells
    (most) and in a neighbour of each, rotations near one rotation (so that
    many pairs pass the rotation test) with some far off."""
    rng = np.random.default_rng(seed)
    a, b = colliding_cells()
    R0 = synth.random_rotation(rng)
    cells = [a, b, a + np.array([1, 0, 0]), b + np.array([0, -1, 0])]
    T = []
    for ci, c in enumerate(cells):
        n = per_cell if ci < 2 else per_cell // 3
        for _ in range(n):
            M = np.eye(4, dtype=np.float32)
            ang = rng.normal(0, 0.05) if rng.uniform() < 0.8 else rng.uniform(0.5, 3.0)
            axis = synth._normalize(rng.normal(size=(1, 3)))[0]
            K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
            Ra = np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * K @ K
            M[:3, :3] = R0 @ Ra
            M[:3, 3] = (c + rng.uniform(0.05, 0.95, 3)) * d_dist
            T.append(M)
    T = np.stack(T)
    perm = rng.permutation(len(T))
    T = T[perm]
    P = len(T)
    counts = rng.integers(1, 40, P).astype(np.uint32)
    codes = (np.uint64(7) << np.uint64(32)) | (rng.integers(0, 50, P).astype(np.uint64) << np.uint64(6)) | \
        np.uint64(5)
    return T, codes, counts, float(d_dist), (a, b)
