"""ctypes binding of the CPU oracle (oracle/libppf_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by
bench.py's cpu_baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libppf_oracle.so")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.oracle_fnv1a.argtypes = [_u8p, C.c_int]
        L.oracle_fnv1a.restype = C.c_uint32
        L.oracle_quant_downf.argtypes = [C.c_float, C.c_float]
        L.oracle_quant_downf.restype = C.c_float
        L.oracle_compute_ppf.argtypes = [_f32p] * 6 + [C.c_float]
        L.oracle_feature_key.argtypes = [_f32p]
        L.oracle_feature_key.restype = C.c_uint32
        L.oracle_cloud_keys.argtypes = [_f32p, _f32p, C.c_int64, C.c_int, C.c_float, _u32p]
        L.oracle_hash_array.argtypes = [_u32p, C.c_int64, _u32p, _u64p, _u64p, _u64p,
                                        C.POINTER(C.c_int64)]
        L.oracle_alpha_literal.argtypes = [_f32p] * 6
        L.oracle_alpha_literal.restype = C.c_uint32
        L.oracle_vote.argtypes = [_f32p, _f32p, C.c_int64, _f32p, _f32p, C.c_int64, C.c_float,
                                  C.c_int, C.c_float, C.c_int, C.c_void_p, C.c_int64, _u64p,
                                  _u32p, C.c_int64, C.POINTER(C.c_int64), _i64p]
        L.oracle_poses.argtypes = [_u64p, C.c_int64, _f32p, _f32p, _f32p, _f32p, _f32p]
        L.oracle_cluster_spatial.argtypes = [_f32p, _u64p, _u32p, C.c_void_p, C.c_int64,
                                             C.c_float, C.c_int, C.c_int, _f32p, _f32p, _f32p,
                                             _u32p, C.POINTER(C.c_int64)]
        L.oracle_cluster_greedy.argtypes = [_f32p, _u32p, C.c_int64, C.c_float, C.c_float,
                                            _f32p, _u32p, C.POINTER(C.c_int64),
                                            C.POINTER(C.c_int64)]
        L.oracle_register.argtypes = [_f32p, _f32p, C.c_int64, _f32p, _f32p, C.c_int64,
                                      C.c_float, C.c_int, C.c_float, C.c_int, C.c_int, C.c_int,
                                      C.c_int, _f32p, _i64p]
        L.oracle_vote_throughput.argtypes = [_f32p, _f32p, C.c_int64, _f32p, _f32p, C.c_int64,
                                             C.c_float, _i64p, C.c_int64, _i64p,
                                             C.POINTER(C.c_double)]
        L.oracle_num_threads.restype = C.c_int
        L.oracle_rot_angle_between.argtypes = [_f32p, _f32p]
        L.oracle_rot_angle_between.restype = C.c_float
        L.oracle_hrotmat2quat.argtypes = [_f32p, _f32p]
        L.oracle_voxel_downsample.argtypes = [_f32p, _f32p, C.c_int64, C.c_float, _f32p, _f32p,
                                              C.POINTER(C.c_int64)]
        L.oracle_numerics.argtypes = [C.c_int, _f32p, _f32p, _f32p, C.c_int64]
        L.oracle_model_create.argtypes = [_f32p, _f32p, C.c_int64, C.c_float]
        L.oracle_model_create.restype = C.c_void_p
        L.oracle_model_destroy.argtypes = [C.c_void_p]
        L.oracle_vote_throughput_h.argtypes = [C.c_void_p, _f32p, _f32p, _f32p, _f32p, C.c_int64,
                                               C.c_float, _i64p, C.c_int64, _i64p,
                                               C.POINTER(C.c_double)]
        _lib = L
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def fnv1a(data: bytes) -> int:
    return int(lib().oracle_fnv1a(np.frombuffer(data, dtype=np.uint8).copy(), len(data)))


def quant_downf(x: float, y: float) -> float:
    return float(lib().oracle_quant_downf(x, y))


def compute_ppf(p1, n1, p2, n2, d_dist):
    raw = np.zeros(4, np.float32)
    q = np.zeros(4, np.float32)
    lib().oracle_compute_ppf(_c(p1, np.float32), _c(n1, np.float32), _c(p2, np.float32),
                             _c(n2, np.float32), raw, q, d_dist)
    return raw, q


def feature_key(f4) -> int:
    return int(lib().oracle_feature_key(_c(f4, np.float32)))


def cloud_keys(points, normals, rpdf: int, d_dist: float) -> np.ndarray:
    n = len(points)
    out = np.zeros(n * n, np.uint32)
    rc = lib().oracle_cloud_keys(_c(points, np.float32), _c(normals, np.float32), n, rpdf,
                                 d_dist, out)
    assert rc == 0
    return out


def hash_array(keys):
    keys = _c(keys, np.uint32)
    n = len(keys)
    hk = np.zeros(max(n, 1), np.uint32)
    cnt = np.zeros(max(n, 1), np.uint64)
    first = np.zeros(max(n, 1), np.uint64)
    k2d = np.zeros(max(n, 1), np.uint64)
    u = C.c_int64()
    lib().oracle_hash_array(keys, n, hk, cnt, first, k2d, C.byref(u))
    u = u.value
    return hk[:u], cnt[:u], first[:u], k2d[:n]


def alpha_literal(m_r, n_r_m, m_i, s_r, n_r_s, s_i) -> int:
    a = [_c(x, np.float32) for x in (m_r, n_r_m, m_i, s_r, n_r_s, s_i)]
    return int(lib().oracle_alpha_literal(*a))


def vote(model, scene, d_dist, rpdf, thr, mode=1, refs=None, cap=1 << 22, per_ref=False):
    """Thresholded, ordered (codes, counts) and stats dict.  per_ref: only
    each reference point's best bin (largest count, then smallest code)."""
    if per_ref:
        mode |= 4
    mp, mn = model
    sp, sn = scene
    codes = np.zeros(cap, np.uint64)
    counts = np.zeros(cap, np.uint32)
    n_out = C.c_int64()
    stats = np.zeros(8, np.int64)
    if refs is not None:
        refs = _c(refs, np.int64)
        rp, nr = refs.ctypes.data, len(refs)
    else:
        rp, nr = None, 0
    rc = lib().oracle_vote(_c(mp, np.float32), _c(mn, np.float32), len(mp), _c(sp, np.float32),
                           _c(sn, np.float32), len(sp), d_dist, rpdf, thr, mode, rp, nr, codes,
                           counts, cap, C.byref(n_out), stats)
    if rc == 2:
        return vote(model, scene, d_dist, rpdf, thr, mode, refs, cap=int(n_out.value) + 1)
    assert rc == 0, rc
    k = n_out.value
    st = dict(lookups=int(stats[0]), hits=int(stats[1]), votes=int(stats[2]),
              unique=int(stats[3]), max_count=int(stats[4]), kept=int(stats[5]))
    return codes[:k].copy(), counts[:k].copy(), st


def poses(codes, model, scene):
    codes = _c(codes, np.uint64)
    P = len(codes)
    T = np.zeros(max(P, 1) * 16, np.float32)
    lib().oracle_poses(codes, P, _c(model[0], np.float32), _c(model[1], np.float32),
                       _c(scene[0], np.float32), _c(scene[1], np.float32), T)
    return T[:P * 16].reshape(P, 4, 4)


def cluster_spatial(T, codes, counts, d_dist, use_l1=False, use_avg=False, weights=None):
    T = _c(T, np.float32).reshape(-1)
    P = len(codes)
    vc = np.zeros(max(P, 1), np.float32)
    tr = np.zeros(max(P, 1) * 3, np.float32)
    q = np.zeros(max(P, 1) * 4, np.float32)
    h = np.zeros(max(P, 1), np.uint32)
    mi = C.c_int64()
    w = _c(weights, np.float32) if weights is not None else None
    lib().oracle_cluster_spatial(T, _c(codes, np.uint64), _c(counts, np.uint32),
                                 w.ctypes.data if w is not None else None, P, d_dist,
                                 int(use_l1), int(use_avg), vc, tr, q, h, C.byref(mi))
    return dict(vote_counts=vc[:P], trans=tr[:3 * P].reshape(P, 3), quats=q[:4 * P].reshape(P, 4),
                hashes=h[:P], max_idx=int(mi.value))


def cluster_greedy(T, counts, trans_thresh, rot_thresh):
    T = _c(T, np.float32).reshape(-1)
    P = len(counts)
    out = np.zeros(3 * 16, np.float32)
    votes = np.zeros(3, np.uint32)
    nc = C.c_int64()
    nr = C.c_int64()
    lib().oracle_cluster_greedy(T, _c(counts, np.uint32), P, trans_thresh, rot_thresh, out, votes,
                                C.byref(nc), C.byref(nr))
    k = nr.value
    return out[:16 * k].reshape(k, 4, 4), votes[:k], int(nc.value)


def rot_angle_between(A, B) -> float:
    """Angle of A^-1 B as the greedy clustering's pose test measures it."""
    return float(lib().oracle_rot_angle_between(_c(A, np.float32).reshape(-1),
                                                _c(B, np.float32).reshape(-1)))


def hrotmat2quat(T) -> np.ndarray:
    """Scalar-first quaternion of a 4x4 as the GPU clustering computes it."""
    q = np.zeros(4, np.float32)
    lib().oracle_hrotmat2quat(_c(T, np.float32).reshape(-1), q)
    return q


def register(model, scene, d_dist, rpdf, thr, cpu_clustering, use_l1=False, use_avg=False,
             per_ref=False):
    pose = np.zeros(16, np.float32)
    stats = np.zeros(8, np.int64)
    rc = lib().oracle_register(_c(model[0], np.float32), _c(model[1], np.float32), len(model[0]),
                               _c(scene[0], np.float32), _c(scene[1], np.float32), len(scene[0]),
                               d_dist, rpdf, thr, int(cpu_clustering), int(use_l1), int(use_avg),
                               int(per_ref), pose, stats)
    st = dict(rc=rc, lookups=int(stats[0]), hits=int(stats[1]), votes=int(stats[2]),
              unique=int(stats[3]), max_count=int(stats[4]), kept=int(stats[5]),
              cluster=int(stats[6]))
    return pose.reshape(4, 4), st


def vote_throughput(model, scene, d_dist, refs):
    refs = _c(refs, np.int64)
    stats = np.zeros(8, np.int64)
    sec = C.c_double()
    rc = lib().oracle_vote_throughput(_c(model[0], np.float32), _c(model[1], np.float32),
                                      len(model[0]), _c(scene[0], np.float32),
                                      _c(scene[1], np.float32), len(scene[0]), d_dist, refs,
                                      len(refs), stats, C.byref(sec))
    assert rc == 0
    return dict(lookups=int(stats[0]), hits=int(stats[1]), votes=int(stats[2]), entries=int(stats[3]),
                seconds=float(sec.value))


def numerics(op: int, a, b=None) -> np.ndarray:
    """Host value of a numerics primitive (op codes of ppf_debug_numerics)."""
    a = _c(a, np.float32)
    n = len(a) // 2 if op == 8 else len(a)
    bb = _c(b if b is not None else np.zeros_like(a), np.float32)
    out = np.zeros(n, np.float32)
    assert lib().oracle_numerics(op, a, bb, out, n) == 0
    return out


class OracleModel:
    """Oracle model hash array built once, for repeated vote timing."""

    def __init__(self, model, d_dist):
        self.model = (_c(model[0], np.float32), _c(model[1], np.float32))
        self.d_dist = float(d_dist)
        self.h = lib().oracle_model_create(self.model[0], self.model[1], len(self.model[0]),
                                           self.d_dist)
        assert self.h, "oracle_model_create failed"

    def vote_throughput(self, scene, refs):
        refs = _c(refs, np.int64)
        stats = np.zeros(8, np.int64)
        sec = C.c_double()
        rc = lib().oracle_vote_throughput_h(self.h, self.model[0], self.model[1],
                                            _c(scene[0], np.float32), _c(scene[1], np.float32),
                                            len(scene[0]), self.d_dist, refs, len(refs), stats,
                                            C.byref(sec))
        assert rc == 0
        return dict(lookups=int(stats[0]), hits=int(stats[1]), votes=int(stats[2]), entries=int(stats[3]),
                    seconds=float(sec.value))

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_model_destroy(self.h)
            self.h = None


def num_threads() -> int:
    return int(lib().oracle_num_threads())


def voxel_downsample(points, normals, leaf: float):
    """pcl::VoxelGrid restatement (alignment.cpp:78-87): (points, normals)."""
    p = _c(np.asarray(points, np.float32).reshape(-1, 3), np.float32)
    n = _c(np.asarray(normals, np.float32).reshape(-1, 3), np.float32)
    op = np.zeros_like(p)
    on = np.zeros_like(n)
    k = C.c_int64()
    assert lib().oracle_voxel_downsample(p, n, len(p), float(leaf), op, on, C.byref(k)) == 0
    return op[:k.value].copy(), on[:k.value].copy()
