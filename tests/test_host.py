"""CPU-side checks of the product library and host logic: the C ABI loads and
exports every function include/ppf_amd.h declares, argument errors come back
as status codes (no exit, no GPU work), and the workload generator is
deterministic."""
import ctypes as C
import hashlib
import json
import os

import numpy as np
import pytest

import ppf_amd
import synth

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_library_exports_declared_symbols():
    names = ppf_amd.exported_symbols_from_header()
    assert "ppf_lookup" in names and "ppf_registration" in names and len(names) >= 20
    lib = C.CDLL(ppf_amd.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_argument_errors_are_status_codes():
    L = ppf_amd.lib()
    h = C.c_void_p()
    pts = np.zeros((4, 3), np.float32)
    # NULL context: rejected before any device work
    rc = L.ppf_model_build(None, pts, pts, 4, 0.1, C.byref(h))
    assert rc == ppf_amd.PPF_ERR_ARG
    assert b"NULL" in L.ppf_last_error()
    rc = L.ppf_scene_upload(None, pts, pts, 4, 1, C.byref(h))
    assert rc == ppf_amd.PPF_ERR_ARG
    rc = L.ppf_lookup(None, None, None, C.byref(h))
    assert rc == ppf_amd.PPF_ERR_ARG
    out = (C.c_void_p * 1)()
    prm = ppf_amd.LookupParams(0.3, 0, 0, 0, 0)
    rc = L.ppf_batch_lookup(None, 1, None, None, C.byref(prm), out, None)
    assert rc == ppf_amd.PPF_ERR_ARG
    # destroy on NULL is a no-op
    assert L.ppf_model_destroy(None) == 0 and L.ppf_result_destroy(None) == 0


def test_python_wrapper_raises():
    with pytest.raises(ppf_amd.PPFError):
        ppf_amd._check(ppf_amd.PPF_ERR_ARG)
    with pytest.raises(ValueError):
        ppf_amd._cloud(np.zeros((3, 3)), np.zeros((2, 3)))


def test_header_limits_consistent():
    text = open(ppf_amd.HEADER).read()
    assert "#define PPF_MAX_MODEL_POINTS 65535" in text
    assert ppf_amd.MAX_MODEL_POINTS == 65535


def _digest(w):
    h = hashlib.sha256()
    for p, n in w.models + w.scenes:
        h.update(p.tobytes())
        h.update(n.tobytes())
    h.update(np.array(w.d_dists, np.float32).tobytes())
    return h.hexdigest()


def test_generator_is_deterministic_and_pinned():
    a = synth.plain_small()
    b = synth.plain_small()
    assert _digest(a) == _digest(b)
    with open(os.path.join(GOLDEN, "workload_digests.json")) as f:
        want = json.load(f)
    assert _digest(a) == want["plain_small"]
    assert len(a.models[0][0]) == 2000 and len(a.scenes[0][0]) == 5000
    assert np.allclose(np.linalg.norm(a.models[0][1], axis=1), 1.0, atol=1e-5)


def test_dataset_like_shape():
    w = synth.dataset_like()
    assert len(w.models) == 4 and len(w.scenes[0][0]) == 50000
    assert all(3000 <= len(m[0]) <= 8000 for m in w.models)
    for m, d in zip(w.models, w.d_dists):
        D = synth.bbox_diameter(m[0])
        assert 0.03 * D * 0.999 <= d <= 0.05 * D * 1.001
