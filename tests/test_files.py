"""PLY / .xf files and the ht_dist validation of the reference driver
(src/alignment.cpp:207-244,300-331; include/impl/util.hpp:95-104;
src/cuda/linalg.cu:9-20).  Host code only (no device)."""
import math
import struct

import numpy as np
import pytest

import ppf_amd


def _rot(axis, ang):
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * K @ K


def _pose(R, t):
    T = np.eye(4)
    T[:3, :3] = R
    T[:3, 3] = t
    return T.astype(np.float32)


@pytest.mark.parametrize("binary", [False, True])
def test_ply_round_trip(tmp_path, binary):
    rng = np.random.default_rng(1)
    p = rng.normal(size=(500, 3)).astype(np.float32)
    n = rng.normal(size=(500, 3)).astype(np.float32)
    f = str(tmp_path / "c.ply")
    ppf_amd.write_ply(f, p, n, binary=binary)
    q, m = ppf_amd.read_ply(f)
    assert q.tobytes() == p.tobytes() and m.tobytes() == n.tobytes()


def test_ply_foreign_layouts(tmp_path):
    # ascii with extra properties, a face list, normal_x naming, CRLF lines
    f = tmp_path / "a.ply"
    f.write_bytes(b"ply\r\nformat ascii 1.0\r\ncomment made by hand\r\nelement vertex 2\r\n"
                  b"property float x\r\nproperty float y\r\nproperty float z\r\nproperty uchar red\r\n"
                  b"property float normal_x\r\nproperty float normal_y\r\nproperty float normal_z\r\n"
                  b"element face 1\r\nproperty list uchar int vertex_indices\r\nend_header\r\n"
                  b"1 2 3 255 0 0 1\r\n4 5 6 7 0 1 0\r\n3 0 1 1\r\n")
    p, n = ppf_amd.read_ply(str(f))
    np.testing.assert_array_equal(p, np.float32([[1, 2, 3], [4, 5, 6]]))
    np.testing.assert_array_equal(n, np.float32([[0, 0, 1], [0, 1, 0]]))
    # big-endian doubles, no normals (read as 0), face element before vertex
    g = tmp_path / "b.ply"
    head = (b"ply\nformat binary_big_endian 1.0\nelement face 1\nproperty list uchar int vertex_indices\n"
            b"element vertex 2\nproperty double x\nproperty double y\nproperty double z\nend_header\n")
    body = struct.pack(">B3i", 3, 0, 1, 0) + struct.pack(">6d", 0.5, -1.25, 2.0, 3.0, 4.5, -6.0)
    g.write_bytes(head + body)
    p, n = ppf_amd.read_ply(str(g))
    np.testing.assert_array_equal(p, np.float32([[0.5, -1.25, 2.0], [3.0, 4.5, -6.0]]))
    assert not n.any()


def test_ply_errors(tmp_path):
    with pytest.raises(ppf_amd.PPFError):
        ppf_amd.read_ply(str(tmp_path / "missing.ply"))
    bad = tmp_path / "bad.ply"
    bad.write_bytes(b"ply\nformat binary_little_endian 1.0\nelement vertex 3\nproperty float x\n"
                    b"property float y\nproperty float z\nend_header\n" + b"\0" * 20)
    with pytest.raises(ppf_amd.PPFError) as e:
        ppf_amd.read_ply(str(bad))
    assert "ends early" in str(e.value)


def test_xf(tmp_path):
    T = _pose(_rot([1, 2, 3], 0.7), [0.1, -2, 3.5])
    f = tmp_path / "gt.xf"
    f.write_text("\n".join(" ".join(f"{v:.9g}" for v in row) for row in T) + "\n")
    np.testing.assert_array_equal(ppf_amd.read_xf(str(f)), T)
    (tmp_path / "short.xf").write_text("1 0 0 0\n0 1 0 0\n")
    with pytest.raises(ppf_amd.PPFError):
        ppf_amd.read_xf(str(tmp_path / "short.xf"))


def test_ht_dist_closed_forms():
    rng = np.random.default_rng(2)
    for _ in range(50):
        R = _rot(rng.normal(size=3), rng.uniform(0, math.pi))
        t = rng.normal(size=3)
        ang = rng.uniform(0, 3.1)
        dt = rng.normal(size=3) * 0.3
        a = _pose(R, t)
        b = _pose(R @ _rot(rng.normal(size=3), ang), t + dt)
        d_t, d_r = ppf_amd.ht_dist(a, b)
        assert abs(d_t - np.linalg.norm(dt)) < 1e-5 * (1 + np.linalg.norm(dt))
        assert abs(d_r - ang) < 2e-4
        # symmetric, and zero against itself
        assert abs(ppf_amd.ht_dist(b, a)[1] - d_r) < 1e-5
        d0 = ppf_amd.ht_dist(a, a)
        assert d0[0] == 0.0 and d0[1] < 1e-6
    # a half turn (trace -1 branch) and the driver's 12 degree threshold
    assert abs(ppf_amd.ht_dist(np.eye(4), _pose(_rot([0, 0, 1], math.pi), [0, 0, 0]))[1] - math.pi) < 1e-5
    assert ppf_amd.ht_dist(np.eye(4), _pose(_rot([1, 1, 0], math.radians(11.9)), [0, 0, 0]))[1] < \
        math.radians(12)
