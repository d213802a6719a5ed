"""Python binding of libppf_amd.so, mirroring the reference's host API.

The reference exposes ``Scene(cloud, d_dist, ref_point_downsample_factor)``
(include/scene.h:15-16), ``Model(cloud, d_dist, vote_count_threshold,
cpu_clustering, use_l1_norm, use_averaged_clusters)`` (include/model.h:17-19)
with ``Model.ppf_lookup(scene)`` (include/model.h:35), and
``ppf_registration(...)`` (include/ppf.h:9-15).  The same names and argument
meanings are kept; clouds are (points[n,3], normals[n,3]) float32 arrays.

Every call goes through the C ABI in include/ppf_amd.h into hand-written HIP
kernels.  There is no CPU fallback: if libppf_amd.so is missing or no GPU is
present, the calls raise.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import weakref
from typing import List, Optional, Sequence, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libppf_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "ppf_amd.h")

PPF_OK = 0
PPF_ERR_ARG = 1
PPF_ERR_HIP = 2
PPF_ERR_NOMEM = 3
PPF_ERR_NO_VOTES = 4
PPF_ERR_STATE = 5
MAX_MODEL_POINTS = 65535


class PPFError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class ModelInfo(C.Structure):
    _fields_ = [("n_points", C.c_int64), ("d_dist", C.c_float), ("n_pairs", C.c_int64),
                ("n_unique_keys", C.c_int64), ("tile_points", C.c_int32), ("n_tiles", C.c_int32),
                ("device_bytes", C.c_uint64)]


class LookupParams(C.Structure):
    _fields_ = [("vote_count_threshold", C.c_float), ("cpu_clustering", C.c_int32),
                ("use_l1_norm", C.c_int32), ("use_averaged_clusters", C.c_int32),
                ("exact_votes", C.c_int32), ("candidate_capacity", C.c_uint32),
                ("per_reference_argmax", C.c_int32)]


class ResultInfo(C.Structure):
    _fields_ = [("status", C.c_int32), ("lookups", C.c_int64), ("votes", C.c_int64),
                ("max_count", C.c_int64),
                ("n_candidates", C.c_int64), ("n_kept", C.c_int64), ("max_idx", C.c_int64),
                ("n_clusters", C.c_int64), ("vote_ms", C.c_double), ("total_ms", C.c_double),
                ("select_ms", C.c_double), ("pose_ms", C.c_double), ("cluster_ms", C.c_double),
                ("cluster_tests", C.c_int64), ("vote_rows_ungrouped", C.c_int64),
                ("cluster_mixed_chunks", C.c_int64), ("cluster_kernel_ms", C.c_double)]


class BatchInfo(C.Structure):
    _fields_ = [("n_jobs", C.c_int64), ("lookups", C.c_int64), ("votes", C.c_int64),
                ("vote_launches", C.c_int32), ("lanes", C.c_int32), ("vote_ms", C.c_double),
                ("finish_ms", C.c_double), ("total_ms", C.c_double), ("fused", C.c_int32),
                ("vote_graph", C.c_int32), ("select_ms", C.c_double), ("pose_ms", C.c_double),
                ("cluster_ms", C.c_double), ("cluster_kernel_ms", C.c_double)]


class TuneInfo(C.Structure):
    _fields_ = [("vote_blocks_per_cu", C.c_int32), ("finish_lanes", C.c_int32),
                ("sort_tile_keys", C.c_int32), ("sort_tile_threads", C.c_int32), ("best_vote_ms", C.c_double),
                ("best_finish_ms", C.c_double), ("best_sort_ms", C.c_double), ("tune_ms", C.c_double)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("max_active", C.c_int32), ("lanes", C.c_int32),
                ("round_ms", C.c_double)]


class RequestInfo(C.Structure):
    _fields_ = [("status", C.c_int32), ("n_targets", C.c_int32), ("vote_rounds", C.c_int32),
                ("pad", C.c_int32), ("queue_ms", C.c_double), ("latency_ms", C.c_double),
                ("device_bytes", C.c_uint64)]


class ServerStats(C.Structure):
    _fields_ = [(k, C.c_int64) for k in ("submitted", "completed", "failed", "cancelled", "rejected",
                                         "rounds", "rows_voted", "queued", "active")] + \
               [(k, C.c_uint64) for k in ("inflight_bytes", "peak_inflight_bytes", "budget_bytes")] + \
               [("max_requests_in_round", C.c_int64), ("vote_ms", C.c_double),
                ("finish_groups", C.c_int64), ("max_finish_group", C.c_int64)]


_lib = None
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_vp = C.c_void_p


_PROGRESS_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_char_p, C.c_int64, C.c_int64)


def lib():
    """Load libppf_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PPFError(PPF_ERR_STATE, f"{LIB_PATH} is missing: run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    L.ppf_last_error.restype = C.c_char_p
    L.ppf_ctx_create.argtypes = [C.c_int, C.c_uint64, C.POINTER(_vp)]
    L.ppf_ctx_destroy.argtypes = [_vp]
    L.ppf_ctx_device_bytes.argtypes = [_vp]
    L.ppf_ctx_device_bytes.restype = C.c_uint64
    L.ppf_ctx_set_progress.argtypes = [_vp, _PROGRESS_FN, C.c_void_p, C.c_int64]
    L.ppf_model_build.argtypes = [_vp, _f32p, _f32p, C.c_int64, C.c_float, C.POINTER(_vp)]
    L.ppf_model_destroy.argtypes = [_vp]
    L.ppf_model_info.argtypes = [_vp, C.POINTER(ModelInfo)]
    L.ppf_model_hash_array.argtypes = [_vp, _vp, _vp, _vp, _vp]
    L.ppf_model_set_weights.argtypes = [_vp, _f32p, C.c_int64]
    L.ppf_scene_upload.argtypes = [_vp, _f32p, _f32p, C.c_int64, C.c_uint32, C.POINTER(_vp)]
    L.ppf_scene_destroy.argtypes = [_vp]
    L.ppf_scene_keys.argtypes = [_vp, C.c_float, _u32p]
    L.ppf_lookup.argtypes = [_vp, _vp, C.POINTER(LookupParams), C.POINTER(_vp)]
    L.ppf_result_destroy.argtypes = [_vp]
    L.ppf_lookup_noalloc.argtypes = [_vp, _vp, C.POINTER(LookupParams), _f32p, C.POINTER(ResultInfo)]
    L.ppf_result_info.argtypes = [_vp, C.POINTER(ResultInfo)]
    L.ppf_result_votes.argtypes = [_vp, _vp, _vp]
    L.ppf_result_transforms.argtypes = [_vp, _f32p]
    L.ppf_result_clusters.argtypes = [_vp, _vp, _vp, _vp]
    L.ppf_result_pose.argtypes = [_vp, _f32p]
    L.ppf_result_cpu_clusters.argtypes = [_vp, _vp, _vp, C.POINTER(C.c_int32)]
    L.ppf_registration.argtypes = [_vp, C.c_int32, _vp, _vp, _vp, C.c_int32, _vp, _vp, _vp, _vp,
                                   C.c_uint32, C.c_float, C.c_int32, C.c_int32, C.c_int32, _vp,
                                   _f32p, _vp]
    L.ppf_model_save.argtypes = [_vp, C.c_char_p]
    L.ppf_model_load.argtypes = [_vp, C.c_char_p, C.POINTER(_vp)]
    L.ppf_voxel_downsample.argtypes = [_vp, _f32p, _f32p, C.c_int64, C.c_float, _vp, _vp,
                                       C.POINTER(C.c_int64)]
    L.ppf_bbox_extent.argtypes = [_f32p, C.c_int64, _f32p]
    L.ppf_ply_read.argtypes = [C.c_char_p, _vp, _vp, C.c_int64, C.POINTER(C.c_int64)]
    L.ppf_ply_write.argtypes = [C.c_char_p, _f32p, _vp, C.c_int64, C.c_int32]
    L.ppf_xf_read.argtypes = [C.c_char_p, _f32p]
    L.ppf_ht_dist.argtypes = [_f32p, _f32p, _f32p]
    L.ppf_batch_lookup.argtypes = [_vp, C.c_int32, _vp, _vp, C.POINTER(LookupParams), _vp,
                                   C.POINTER(BatchInfo)]
    L.ppf_batch_tune.argtypes = [_vp, C.c_int32, _vp, _vp, C.POINTER(LookupParams), C.POINTER(TuneInfo)]
    L.ppf_tune_reset.argtypes = [_vp]
    L.ppf_server_create.argtypes = [_vp, C.c_int32, _vp, C.POINTER(ServerConfig), C.POINTER(_vp)]
    L.ppf_server_destroy.argtypes = [_vp]
    L.ppf_server_submit.argtypes = [_vp, _f32p, _f32p, C.c_int64, C.c_uint32, C.c_int32, _vp,
                                    C.POINTER(LookupParams), C.POINTER(C.c_uint64)]
    L.ppf_server_wait.argtypes = [_vp, C.c_uint64, C.c_double, _vp, _vp, _vp, C.POINTER(RequestInfo)]
    L.ppf_server_poll.argtypes = [_vp, C.c_uint64, C.POINTER(C.c_int32)]
    L.ppf_server_cancel.argtypes = [_vp, C.c_uint64]
    L.ppf_server_stats.argtypes = [_vp, C.POINTER(ServerStats)]
    L.ppf_debug_numerics.argtypes = [_vp, C.c_int32, _f32p, _vp, _f32p, C.c_int64]
    L.ppf_debug_primitive.argtypes = [_vp, C.c_int32, _vp, _vp, C.c_int64, C.c_int32, C.c_int32]
    L.ppf_debug_set.argtypes = [_vp, C.c_int32, C.c_int64]
    L.ppf_debug_cluster.argtypes = [_vp, _f32p, _vp, _vp, _vp, C.c_int64, C.c_int64, C.c_float,
                                    C.c_int32, C.c_int32, _f32p, _f32p, _f32p, _vp]
    sizes = (C.c_uint32 * 8)()
    L.ppf_abi_sizes(sizes)
    want = tuple(C.sizeof(t) for t in (ModelInfo, LookupParams, ResultInfo, BatchInfo, ServerConfig,
                                       RequestInfo, ServerStats, TuneInfo))
    if tuple(sizes) != want:
        raise PPFError(PPF_ERR_STATE, f"{LIB_PATH} was built from a different include/ppf_amd.h "
                                      f"(struct sizes {tuple(sizes)} != {want}): rebuild it")
    _lib = L
    return L


def _check(rc: int):
    if rc != PPF_OK:
        raise PPFError(rc, lib().ppf_last_error().decode())


def _cloud(points, normals) -> Tuple[np.ndarray, np.ndarray]:
    p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    n = np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 3)
    if p.shape != n.shape:
        raise ValueError("points and normals must have the same shape")
    return p, n


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data


_live_contexts: "weakref.WeakSet[Context]" = weakref.WeakSet()


class _Owned:
    """Base of objects created from a Context: destroyed before their context."""

    handle = None
    _destroy = None

    def _register(self, ctx: "Context"):
        self.ctx = ctx
        ctx._children.add(self)

    def close(self):
        h = getattr(self, "handle", None)
        if h:
            self.handle = None
            if self.ctx.handle:
                getattr(lib(), self._destroy)(h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """One device + one HIP stream (the reference's device selection, ppf.cu:37-49).

    Closing a context first closes every scene, model and result made from it."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0):
        h = _vp()
        _check(lib().ppf_ctx_create(device, hbm_budget_bytes, C.byref(h)))
        self.handle = h
        self._children: "weakref.WeakSet[_Owned]" = weakref.WeakSet()
        _live_contexts.add(self)

    def close(self):
        if getattr(self, "handle", None):
            # results before models/scenes (they are independent on the device,
            # but keep a deterministic order)
            order = {"Server": 0, "LookupResult": 1}
            kids = sorted(list(self._children), key=lambda o: order.get(type(o).__name__, 2))
            for k in kids:
                k.close()
            lib().ppf_ctx_destroy(self.handle)
            self.handle = None

    def device_bytes(self) -> int:
        return int(lib().ppf_ctx_device_bytes(self.handle))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def debug_numerics(self, op: int, a: np.ndarray, b: Optional[np.ndarray] = None) -> np.ndarray:
        a = np.ascontiguousarray(a, dtype=np.float32)
        n = len(a) // 2 if op == 8 else len(a)
        out = np.zeros(n, np.float32)
        bb = None if b is None else np.ascontiguousarray(b, dtype=np.float32)
        _check(lib().ppf_debug_numerics(self.handle, op, a, _ptr(bb), out, n))
        return out


    def debug_scan(self, a: np.ndarray):
        x = np.ascontiguousarray(a, dtype=np.uint32).copy()
        tot = np.zeros(1, np.uint32)
        _check(lib().ppf_debug_primitive(self.handle, 0, x.ctypes.data, tot.ctypes.data, len(x), 0, 0))
        return x, int(tot[0])

    def debug_copy(self, a: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(a, dtype=np.uint32)
        out = np.zeros_like(x)
        _check(lib().ppf_debug_primitive(self.handle, 2, x.ctypes.data, out.ctypes.data, len(x), 0, 0))
        return out

    def debug_fma_stream(self, blocks: int, iters: int) -> np.ndarray:
        """Full-rate VALU reference kernel (ppf_debug_primitive op 3)."""
        out = np.zeros(blocks * 256, np.float32)
        _check(lib().ppf_debug_primitive(self.handle, 3, out.ctypes.data, None, blocks, iters, 0))
        return out

    def set_progress(self, fn=None, rows_per_launch: int = 0):
        """ppf_ctx_set_progress: fn(stage, done, total) after each vote launch
        of rows_per_launch reference rows and after each later stage of a
        plain lookup; None turns reports off."""
        if fn is None:
            self._progress = None
            _check(lib().ppf_ctx_set_progress(self.handle, _PROGRESS_FN(), None, 0))
            return
        cb = _PROGRESS_FN(lambda _u, stage, done, total: fn(stage.decode(), int(done), int(total)))
        self._progress = cb  # keep the C callback alive
        _check(lib().ppf_ctx_set_progress(self.handle, cb, None, int(rows_per_launch)))

    def debug_set(self, knob: int, value: int):
        """ppf_debug_set: knob 0 = distinct buckets a reference row may group
        before it is voted hit by hit (0 = the grouping table's size)."""
        _check(lib().ppf_debug_set(self.handle, int(knob), int(value)))

    def debug_cluster(self, T, codes, counts, d_dist: float, use_l1: bool = False,
                      use_avg: bool = False, weights=None):
        """GPU spatial-hash clustering of host poses: dict of vote_counts,
        trans, quats, max_idx, tests, mixed_chunks."""
        T = np.ascontiguousarray(T, np.float32).reshape(-1)
        codes = np.ascontiguousarray(codes, np.uint64)
        counts = np.ascontiguousarray(counts, np.uint32)
        P = len(codes)
        w = None if weights is None else np.ascontiguousarray(weights, np.float32)
        vc = np.zeros(P, np.float32)
        tr = np.zeros(3 * P, np.float32)
        q = np.zeros(4 * P, np.float32)
        st = np.zeros(3, np.int64)
        _check(lib().ppf_debug_cluster(self.handle, T, codes.ctypes.data, counts.ctypes.data,
                                       None if w is None else w.ctypes.data,
                                       0 if w is None else len(w), P, float(d_dist), int(use_l1),
                                       int(use_avg), vc, tr, q, st.ctypes.data))
        return dict(vote_counts=vc, trans=tr.reshape(P, 3), quats=q.reshape(P, 4),
                    max_idx=int(st[0]), tests=int(st[1]), mixed_chunks=int(st[2]))

    def debug_sort(self, keys: np.ndarray, vals: Optional[np.ndarray], bit_lo: int, bit_hi: int):
        k = np.ascontiguousarray(keys, dtype=np.uint64).copy()
        v = None if vals is None else np.ascontiguousarray(vals, dtype=np.uint32).copy()
        _check(lib().ppf_debug_primitive(self.handle, 1, k.ctypes.data, _ptr(v), len(k), bit_lo, bit_hi))
        return k, v

    def debug_key_sort(self, words: np.ndarray):
        """The model build's stable sort of pair words by their high 32 bits
        (ppf_debug_primitive op 4); returns (sorted words, True when the
        rank-compressed path sorted them)."""
        k = np.ascontiguousarray(words, dtype=np.uint64).copy()
        flag = np.zeros(1, np.uint32)
        _check(lib().ppf_debug_primitive(self.handle, 4, k.ctypes.data, flag.ctypes.data, len(k), 0, 0))
        return k, bool(flag[0])


_default_ctx: Optional[Context] = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class Scene(_Owned):
    """Scene(cloud, d_dist, ref_point_downsample_factor=1) -- include/scene.h:15-16.

    d_dist is accepted for signature parity; scene features are quantized with
    the model's d_dist at lookup time (the reference rebuilds the scene per
    model for the same reason, src/cuda/ppf.cu:64-67)."""

    def __init__(self, cloud, d_dist: float = 0.0, ref_point_downsample_factor: int = 1,
                 ctx: Optional[Context] = None):
        ctx = ctx or default_context()
        p, n = _cloud(*cloud)
        self.n = len(p)
        self.d_dist = d_dist
        self.rpdf = int(ref_point_downsample_factor)
        h = _vp()
        _check(lib().ppf_scene_upload(ctx.handle, p, n, self.n, self.rpdf, C.byref(h)))
        self.handle = h
        self._destroy = "ppf_scene_destroy"
        self._register(ctx)

    def numPoints(self) -> int:
        return self.n

    def getHashKeys(self, d_dist: Optional[float] = None) -> np.ndarray:
        """All n*n keys [s_r][s_i] for d_dist (Scene::getHashKeys)."""
        out = np.zeros(self.n * self.n, np.uint32)
        _check(lib().ppf_scene_keys(self.handle, float(d_dist if d_dist else self.d_dist), out))
        return out



class LookupResult(_Owned):
    """Outputs of one Model::ppf_lookup; owns its device arrays."""

    def __init__(self, handle, cpu: bool, ctx: "Context"):
        self.handle = handle
        self._destroy = "ppf_result_destroy"
        self._register(ctx)
        self.cpu = cpu
        info = ResultInfo()
        _check(lib().ppf_result_info(handle, C.byref(info)))
        self.info = {k: getattr(info, k) for k, _ in ResultInfo._fields_}

    @property
    def n_kept(self) -> int:
        return int(self.info["n_kept"])

    def votes(self) -> Tuple[np.ndarray, np.ndarray]:
        P = self.n_kept
        codes = np.zeros(max(P, 1), np.uint64)
        counts = np.zeros(max(P, 1), np.uint32)
        _check(lib().ppf_result_votes(self.handle, codes.ctypes.data, counts.ctypes.data))
        return codes[:P], counts[:P]

    def transformations(self) -> np.ndarray:
        P = self.n_kept
        T = np.zeros(max(P, 1) * 16, np.float32)
        _check(lib().ppf_result_transforms(self.handle, T))
        return T[:16 * P].reshape(P, 4, 4)

    def clusters(self):
        P = self.n_kept
        vc = np.zeros(max(P, 1), np.float32)
        tr = np.zeros(max(P, 1) * 3, np.float32)
        q = np.zeros(max(P, 1) * 4, np.float32)
        _check(lib().ppf_result_clusters(self.handle, vc.ctypes.data, tr.ctypes.data, q.ctypes.data))
        return vc[:P], tr[:3 * P].reshape(P, 3), q[:4 * P].reshape(P, 4)

    def cpu_transformations(self):
        """(poses [k][4][4], votes [k]) of the k <= 3 best CPU clusters
        (Model::cpu_transformations, transformation_clustering.cpp:96-122)."""
        poses = np.zeros(3 * 16, np.float32)
        votes = np.zeros(3, np.uint32)
        n = C.c_int32()
        _check(lib().ppf_result_cpu_clusters(self.handle, poses.ctypes.data, votes.ctypes.data,
                                             C.byref(n)))
        k = int(n.value)
        return poses[:16 * k].reshape(k, 4, 4), votes[:k]

    def pose(self) -> np.ndarray:
        out = np.zeros(16, np.float32)
        rc = lib().ppf_result_pose(self.handle, out)
        if rc not in (PPF_OK, PPF_ERR_NO_VOTES):
            _check(rc)
        return out.reshape(4, 4)



class Model(_Owned):
    """Model(cloud, d_dist, vote_count_threshold, cpu_clustering, use_l1_norm,
    use_averaged_clusters) -- include/model.h:17-19, src/cuda/model.cu:43-82."""

    def __init__(self, cloud, d_dist: float, vote_count_threshold: float = 0.4,
                 cpu_clustering: bool = False, use_l1_norm: bool = False,
                 use_averaged_clusters: bool = False, ctx: Optional[Context] = None):
        ctx = ctx or default_context()
        p, n = _cloud(*cloud)
        self.n = len(p)
        self.d_dist = float(d_dist)
        self.vote_count_threshold = float(vote_count_threshold)
        self.cpu_clustering = bool(cpu_clustering)
        self.use_l1_norm = bool(use_l1_norm)
        self.use_averaged_clusters = bool(use_averaged_clusters)
        h = _vp()
        _check(lib().ppf_model_build(ctx.handle, p, n, self.n, self.d_dist, C.byref(h)))
        self.handle = h
        self._destroy = "ppf_model_destroy"
        self._register(ctx)
        self.result: Optional[LookupResult] = None

    def save(self, path: str):
        """Write the built model (device arrays, versioned, checksummed)."""
        _check(lib().ppf_model_save(self.handle, os.fsencode(path)))

    @classmethod
    def load(cls, path: str, vote_count_threshold: float = 0.4, cpu_clustering: bool = False,
             use_l1_norm: bool = False, use_averaged_clusters: bool = False,
             ctx: Optional[Context] = None) -> "Model":
        """A model saved by save(), without rebuilding its hash array."""
        ctx = ctx or default_context()
        h = _vp()
        _check(lib().ppf_model_load(ctx.handle, os.fsencode(path), C.byref(h)))
        self = cls.__new__(cls)
        self.handle = h
        self._destroy = "ppf_model_destroy"
        self._register(ctx)
        self.vote_count_threshold = float(vote_count_threshold)
        self.cpu_clustering = bool(cpu_clustering)
        self.use_l1_norm = bool(use_l1_norm)
        self.use_averaged_clusters = bool(use_averaged_clusters)
        self.result = None
        inf = self.info()
        self.n = int(inf["n_points"])
        self.d_dist = float(inf["d_dist"])
        return self

    def info(self) -> dict:
        i = ModelInfo()
        _check(lib().ppf_model_info(self.handle, C.byref(i)))
        return {k: getattr(i, k) for k, _ in ModelInfo._fields_}

    def hash_array(self, with_pair_map: bool = True):
        """(hashkeys, counts, firstHashkeyIndex, hashkeyToDataMap) --
        ParallelHashArray getters (include/impl/parallel_hash_array.hpp:94-112).
        with_pair_map=False skips the n*n pair map (returned as None)."""
        U = int(self.info()["n_unique_keys"])
        hk = np.zeros(U, np.uint32)
        cnt = np.zeros(U, np.uint32)
        first = np.zeros(U, np.uint32)
        k2d = np.zeros(self.n * self.n, np.uint32) if with_pair_map else None
        _check(lib().ppf_model_hash_array(self.handle, hk.ctypes.data, cnt.ctypes.data,
                                          first.ctypes.data, _ptr(k2d)))
        return hk, cnt, first, k2d

    def SetModelPointVoteWeights(self, weights):
        w = np.ascontiguousarray(weights, dtype=np.float32)
        _check(lib().ppf_model_set_weights(self.handle, w, len(w)))

    def ppf_lookup(self, scene: Scene, exact_votes: bool = False,
                   candidate_capacity: int = 0, per_reference_argmax: bool = False) -> LookupResult:
        prm = LookupParams(self.vote_count_threshold, int(self.cpu_clustering),
                           int(self.use_l1_norm), int(self.use_averaged_clusters),
                           int(exact_votes), int(candidate_capacity), int(per_reference_argmax))
        h = _vp()
        _check(lib().ppf_lookup(self.handle, scene.handle, C.byref(prm), C.byref(h)))
        self.result = LookupResult(h, self.cpu_clustering, self.ctx)
        return self.result

    def ppf_lookup_noalloc(self, scene: Scene, exact_votes: bool = False,
                           per_reference_argmax: bool = False):
        """Allocation-free lookup (ppf_lookup_noalloc): (pose 4x4, info dict).
        Scenes of at most NOALLOC_MAX_SCENE_POINTS points; the device scratch
        is one static array per GPU shared by all such calls."""
        prm = LookupParams(self.vote_count_threshold, int(self.cpu_clustering),
                           int(self.use_l1_norm), int(self.use_averaged_clusters),
                           int(exact_votes), 0, int(per_reference_argmax))
        pose = np.zeros(16, np.float32)
        info = ResultInfo()
        rc = lib().ppf_lookup_noalloc(self.handle, scene.handle, C.byref(prm), pose, C.byref(info))
        if rc not in (PPF_OK, PPF_ERR_NO_VOTES):
            _check(rc)
        return pose.reshape(4, 4), {k: getattr(info, k) for k, _ in ResultInfo._fields_}

    # reference-style accessors of the last lookup
    @property
    def max_idx(self) -> int:
        return int(self.result.info["max_idx"]) if self.result else -1

    def getTransformations(self) -> np.ndarray:
        return self.result.transformations()



def voxel_downsample(cloud, leaf: float, ctx: Optional[Context] = None):
    """pcl::VoxelGrid downsampling with leaf (leaf, leaf, leaf) on the device
    (the reference driver's voxelGridDownsample, src/alignment.cpp:78-87):
    returns (points, normals) of the occupied voxels' averages."""
    ctx = ctx or default_context()
    p, n = _cloud(*cloud)
    op = np.zeros(max(len(p), 1) * 3, np.float32)
    on = np.zeros(max(len(p), 1) * 3, np.float32)
    k = C.c_int64()
    _check(lib().ppf_voxel_downsample(ctx.handle, p, n, len(p), float(leaf), op.ctypes.data,
                                      on.ctypes.data, C.byref(k)))
    return op[:3 * k.value].reshape(-1, 3).copy(), on[:3 * k.value].reshape(-1, 3).copy()


def bbox_extent(points) -> np.ndarray:
    """max - min per axis over finite points (pcl::getMinMax3D, alignment.cpp:246-257)."""
    p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    out = np.zeros(3, np.float32)
    _check(lib().ppf_bbox_extent(p, len(p), out))
    return out


def read_ply(path: str):
    """(points[n,3], normals[n,3]) of a PLY file (pcl::io::loadPLYFile into
    PointNormal, src/alignment.cpp:207-244)."""
    b = os.fsencode(path)
    n = C.c_int64()
    _check(lib().ppf_ply_read(b, None, None, 0, C.byref(n)))
    p = np.zeros((n.value, 3), np.float32)
    nn = np.zeros((n.value, 3), np.float32)
    _check(lib().ppf_ply_read(b, p.ctypes.data, nn.ctypes.data, n.value, C.byref(n)))
    return p, nn


def write_ply(path: str, points, normals=None, binary: bool = True):
    p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    nn = None if normals is None else np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 3)
    _check(lib().ppf_ply_write(os.fsencode(path), p, _ptr(nn), len(p), int(binary)))


def read_xf(path: str) -> np.ndarray:
    """4x4 ground-truth pose of a .xf file (include/impl/util.hpp:95-104)."""
    out = np.zeros(16, np.float32)
    _check(lib().ppf_xf_read(os.fsencode(path), out))
    return out.reshape(4, 4)


def ht_dist(a, b):
    """(translation distance, rotation angle in rad) between two poses
    (src/cuda/linalg.cu:9-20)."""
    out = np.zeros(2, np.float32)
    _check(lib().ppf_ht_dist(np.ascontiguousarray(a, np.float32).reshape(16),
                             np.ascontiguousarray(b, np.float32).reshape(16), out))
    return float(out[0]), float(out[1])


def batch_lookup(models: Sequence["Model"], scenes: Sequence[Scene], vote_count_threshold: float,
                 cpu_clustering: bool = False, use_l1_norm: bool = False,
                 use_averaged_clusters: bool = False, exact_votes: bool = False,
                 ctx: Optional[Context] = None, candidate_capacity: int = 0,
                 per_reference_argmax: bool = False):
    """Batched form: result j of (models[j], scenes[j]), all jobs voted in one
    persistent launch (ppf_batch_lookup).  Returns (results, info dict)."""
    if len(models) != len(scenes) or not models:
        raise ValueError("batch_lookup needs one scene per model and at least one job")
    ctx = ctx or models[0].ctx
    n = len(models)
    mh = (C.c_void_p * n)(*[m.handle.value for m in models])
    sh = (C.c_void_p * n)(*[s.handle.value for s in scenes])
    prm = LookupParams(float(vote_count_threshold), int(cpu_clustering), int(use_l1_norm),
                       int(use_averaged_clusters), int(exact_votes), int(candidate_capacity),
                       int(per_reference_argmax))
    out = (C.c_void_p * n)()
    info = BatchInfo()
    _check(lib().ppf_batch_lookup(ctx.handle, n, mh, sh, C.byref(prm), out, C.byref(info)))
    res = [LookupResult(_vp(out[j]), bool(cpu_clustering), ctx) for j in range(n)]
    return res, {k: getattr(info, k) for k, _ in BatchInfo._fields_}


def batch_tune(models: Sequence["Model"], scenes: Sequence[Scene], vote_count_threshold: float,
               ctx: Optional[Context] = None) -> dict:
    """Autotuner of the batched form (ppf_batch_tune): times launch shapes on
    this batch and keeps the fastest on the context; returns the choice."""
    ctx = ctx or models[0].ctx
    n = len(models)
    mh = (C.c_void_p * n)(*[m.handle.value for m in models])
    sh = (C.c_void_p * n)(*[s.handle.value for s in scenes])
    prm = LookupParams(float(vote_count_threshold), 0, 0, 0, 0, 0, 0)
    info = TuneInfo()
    _check(lib().ppf_batch_tune(ctx.handle, n, mh, sh, C.byref(prm), C.byref(info)))
    return {k: getattr(info, k) for k, _ in TuneInfo._fields_}


def tune_reset(ctx: Context):
    _check(lib().ppf_tune_reset(ctx.handle))


NOALLOC_MAX_SCENE_POINTS = 8192
PPF_ERR_TIMEOUT = 6
PPF_ERR_CANCELLED = 7


class Server(_Owned):
    """Served form: a request loop over resident models (ppf_server_*).

    submit() copies the scene and returns a ticket; wait() returns
    (poses[n_targets,4,4], statuses[n_targets], info) and, with
    results=True, the full LookupResult of every target."""

    def __init__(self, models: Sequence["Model"], hbm_budget_bytes: int = 0, max_active: int = 0,
                 lanes: int = 0, round_ms: float = 0.0, ctx: Optional[Context] = None):
        ctx = ctx or models[0].ctx
        self.models = list(models)
        mh = (C.c_void_p * len(models))(*[m.handle.value for m in models])
        cfg = ServerConfig(int(hbm_budget_bytes), int(max_active), int(lanes), float(round_ms))
        h = _vp()
        _check(lib().ppf_server_create(ctx.handle, len(models), mh, C.byref(cfg), C.byref(h)))
        self.handle = h
        self._destroy = "ppf_server_destroy"
        self._register(ctx)
        self._targets = {}

    def submit(self, cloud, model_ids: Sequence[int], ref_point_downsample_factor: int,
               vote_count_threshold: float, cpu_clustering: bool = False, use_l1_norm: bool = False,
               use_averaged_clusters: bool = False, exact_votes: bool = False,
               candidate_capacity: int = 0, per_reference_argmax: bool = False) -> int:
        p, n = _cloud(*cloud)
        ids = np.ascontiguousarray(model_ids, dtype=np.int32)
        prm = LookupParams(float(vote_count_threshold), int(cpu_clustering), int(use_l1_norm),
                           int(use_averaged_clusters), int(exact_votes), int(candidate_capacity),
                           int(per_reference_argmax))
        t = C.c_uint64()
        _check(lib().ppf_server_submit(self.handle, p, n, len(p), int(ref_point_downsample_factor),
                                       len(ids), ids.ctypes.data, C.byref(prm), C.byref(t)))
        self._targets[t.value] = (len(ids), bool(cpu_clustering))
        return t.value

    def wait(self, ticket: int, timeout_s: float = -1.0, results: bool = False):
        nt, cpu = self._targets[ticket]
        poses = np.zeros(nt * 16, np.float32)
        st = np.zeros(nt, np.int32)
        res = (C.c_void_p * nt)() if results else None
        info = RequestInfo()
        rc = lib().ppf_server_wait(self.handle, ticket, float(timeout_s), poses.ctypes.data,
                                   st.ctypes.data, res, C.byref(info))
        if rc == PPF_ERR_TIMEOUT:
            raise PPFError(rc, lib().ppf_last_error().decode())
        del self._targets[ticket]
        out_res = None
        if results:
            out_res = [LookupResult(_vp(res[k]), cpu, self.ctx) if res[k] else None for k in range(nt)]
        d = {k: getattr(info, k) for k, _ in RequestInfo._fields_}
        d["rc"] = rc
        if rc != PPF_OK:
            d["error"] = lib().ppf_last_error().decode()
        return poses.reshape(nt, 4, 4), st, d, out_res

    def poll(self, ticket: int) -> bool:
        d = C.c_int32()
        _check(lib().ppf_server_poll(self.handle, ticket, C.byref(d)))
        return bool(d.value)

    def cancel(self, ticket: int) -> int:
        return int(lib().ppf_server_cancel(self.handle, ticket))

    def stats(self) -> dict:
        st = ServerStats()
        _check(lib().ppf_server_stats(self.handle, C.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}


def _close_all_contexts():
    for c in list(_live_contexts):
        try:
            c.close()
        except Exception:
            pass


# tear down before the HIP runtime's own exit-time destructors run
atexit.register(_close_all_contexts)


def ppf_registration(scene_clouds: Sequence, model_clouds: Sequence, model_d_dists: Sequence[float],
                     ref_point_downsample_factor: int, vote_count_threshold: float,
                     cpu_clustering: bool, use_l1_norm: bool = False,
                     use_averaged_clusters: bool = False, devUse: int = 0,
                     model_weights: Optional[Sequence] = None, ctx: Optional[Context] = None,
                     return_status: bool = False):
    """Pose of every model in every scene: list[scene][model] of 4x4 float32
    (include/ppf.h:9-15, src/cuda/ppf.cu:29-106)."""
    ctx = ctx or Context(devUse)
    scenes = [_cloud(*c) for c in scene_clouds]
    models = [_cloud(*c) for c in model_clouds]
    ns, nm = len(scenes), len(models)
    sp = (C.c_void_p * max(ns, 1))(*[s[0].ctypes.data for s in scenes])
    sn = (C.c_void_p * max(ns, 1))(*[s[1].ctypes.data for s in scenes])
    ss = np.array([len(s[0]) for s in scenes] or [0], np.int64)
    mp = (C.c_void_p * max(nm, 1))(*[m[0].ctypes.data for m in models])
    mn = (C.c_void_p * max(nm, 1))(*[m[1].ctypes.data for m in models])
    ms = np.array([len(m[0]) for m in models] or [0], np.int64)
    dd = np.ascontiguousarray(model_d_dists, dtype=np.float32)
    wts = None
    keep = []
    if model_weights is not None:
        arrs = [None if w is None else np.ascontiguousarray(w, dtype=np.float32) for w in model_weights]
        keep = arrs
        wts = (C.c_void_p * max(nm, 1))(*[0 if a is None else a.ctypes.data for a in arrs])
    poses = np.zeros(max(ns * nm, 1) * 16, np.float32)
    status = np.zeros(max(ns * nm, 1), np.int32)
    _check(lib().ppf_registration(ctx.handle, ns, sp, sn, ss.ctypes.data, nm, mp, mn, ms.ctypes.data,
                                  dd.ctypes.data, int(ref_point_downsample_factor),
                                  float(vote_count_threshold), int(cpu_clustering), int(use_l1_norm),
                                  int(use_averaged_clusters), wts, poses, status.ctypes.data))
    del keep
    P = poses[:ns * nm * 16].reshape(ns, nm, 4, 4)
    out = [[P[i, j].copy() for j in range(nm)] for i in range(ns)]
    if return_status:
        return out, status[:ns * nm].reshape(ns, nm)
    return out


def exported_symbols_from_header(path: str = HEADER) -> List[str]:
    """Names of the functions declared in include/ppf_amd.h."""
    import re
    text = open(path).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(ppf_[a-z_0-9]+)\s*\(", text,
                                 flags=re.M)))
