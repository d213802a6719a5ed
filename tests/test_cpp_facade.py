"""The C++ facade (include/ppf_amd.hpp) used the way the reference's host code
uses Scene/Model/ppf_registration (include/scene.h:15-16, include/model.h:17-113,
include/ppf.h:9-15): tests/cpp/facade_main.cpp holds src/cuda/ppf.cu:63-93's
extraction of the pose from Model's public results and alignment.cpp:290-298's
ppf_registration call, changed only in includes and type names.  Compiled and
linked against libppf_amd.so on the CPU, run on the GPU and compared bit for
bit with the Python binding's results."""
import os
import subprocess

import numpy as np
import pytest

from conftest import PKG, ROOT

SRC = os.path.join(ROOT, "tests", "cpp", "facade_main.cpp")
LIB = os.path.join(PKG, "libppf_amd.so")


@pytest.fixture(scope="module")
def facade_bin(tmp_path_factory):
    if not os.path.exists(LIB):
        pytest.fail("libppf_amd.so is not built")
    out = str(tmp_path_factory.mktemp("facade") / "facade_main")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", "-I" + os.path.join(ROOT, "include"),
           SRC, "-o", out, "-L" + PKG, "-lppf_amd", "-Wl,-rpath," + PKG,
           "-Wl,-rpath-link,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return out


def test_facade_compiles_and_links(facade_bin):
    assert os.access(facade_bin, os.X_OK)


def test_header_is_plain_c():
    """include/ppf_amd.h compiles as C99 (no C++ types at the boundary)."""
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-pedantic", "-fsyntax-only", "-x", "c",
                        os.path.join(ROOT, "include", "ppf_amd.h")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("cpu_clustering", [False, True])
def test_facade_matches_python_binding(facade_bin, tmp_path, cpu_clustering):
    import ppf_amd
    import synth
    w = synth.plain_small(seed=11, n_model=400, n_clutter=600)
    m, s, d = w.models[0], w.scenes[0], w.d_dists[0]
    mf, sf, of = tmp_path / "m.f32", tmp_path / "s.f32", tmp_path / "out.bin"
    np.hstack([m[0], m[1]]).astype(np.float32).tofile(mf)
    np.hstack([s[0], s[1]]).astype(np.float32).tofile(sf)
    r = subprocess.run([facade_bin, str(mf), str(sf), repr(float(np.float32(d))), str(w.rpdf),
                        repr(float(np.float32(w.vote_count_threshold))), str(int(cpu_clustering)), str(of)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = of.read_bytes()
    t1 = np.frombuffer(raw[:64], np.float32).reshape(4, 4)
    t2 = np.frombuffer(raw[64:128], np.float32).reshape(4, 4)
    n_kept, max_idx, votes, n_cpu = np.frombuffer(raw[128:160], np.int64)
    w0 = np.frombuffer(raw[160:164], np.float32)[0]

    ctx = ppf_amd.Context(0)
    try:
        model = ppf_amd.Model(m, d, w.vote_count_threshold, cpu_clustering=cpu_clustering, ctx=ctx)
        scene = ppf_amd.Scene(s, d, w.rpdf, ctx=ctx)
        res = model.ppf_lookup(scene)
        ref = res.pose()
        assert res.info["n_kept"] == n_kept and res.info["votes"] == votes
        codes, counts = res.votes()
        assert w0 == np.float32(counts[0])  # weights default to 1 (model.cu:67)
        if not cpu_clustering:
            assert res.info["max_idx"] == max_idx and n_cpu == 0
        else:
            assert 1 <= n_cpu <= 3
    finally:
        ctx.close()
    assert t1.tobytes() == ref.astype(np.float32).tobytes()
    assert t2.tobytes() == t1.tobytes()
