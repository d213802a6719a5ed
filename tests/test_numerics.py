"""CPU pins of the scalar numerics: FNV-1a, quant_downf, features, and the
build's own acos/atan2/sin/cos (oracle/ppf_oracle.cpp, csrc/ppf_math.h).

The reference has no tests or golden vectors; these pins come from textbook
known answers, closed forms and double-precision libm.
"""
import math

import numpy as np
import pytest

import ppf_oracle as O

D = np.float32(np.float32(2.0 * np.float32(3.141592654)) / np.float32(30.0))


def test_angle_step_constant():
    # include/kernel.h:16 with CUDART_PI_F = 3.141592654f
    assert D == np.float32(0.20943951606750488)
    assert np.float32(30) * D == np.float32(2) * np.float32(3.141592654)


@pytest.mark.parametrize("data,expect", [
    (b"", 0x811C9DC5),
    (b"a", 0xE40C292C),
    (b"foobar", 0xBF9CF968),
])
def test_fnv1a_textbook(data, expect):
    # published FNV-1a 32-bit test vectors (ASCII bytes: no sign extension)
    assert O.fnv1a(data) == expect


def test_fnv1a_signed_char_bytes():
    # kernel.cu:24-27 reads bytes through `char *`: 0x80..0xff are sign-extended,
    # so the XOR flips the top 24 bits as well.  Hand derivation for one byte:
    h = (0x811C9DC5 ^ 0xFFFFFF80) * 16777619 % 2**32
    assert O.fnv1a(b"\x80") == h
    unsigned = (0x811C9DC5 ^ 0x80) * 16777619 % 2**32
    assert O.fnv1a(b"\x80") != unsigned
    # bytes of 1.0f (00 00 80 3f): only the third byte is >= 0x80
    h = 0x811C9DC5
    for b, ext in ((0x00, 0x00), (0x00, 0x00), (0x80, 0xFFFFFF80), (0x3F, 0x3F)):
        h = ((h ^ ext) * 16777619) % 2**32
    assert O.fnv1a(np.float32(1.0).tobytes()) == h


def test_feature_key_nan_rules():
    # .x NaN -> key 0 (kernel.cu:467-469); NaN elsewhere hashed as 0x7fc00000
    assert O.feature_key(np.array([np.nan, 1, 2, 3], np.float32)) == 0
    a = np.array([1, np.nan, 2, 3], np.float32)
    b = a.copy().view(np.uint32)
    b[1] = 0xFFC00000  # negative quiet NaN (x86 0/0)
    assert O.feature_key(a) == O.feature_key(b.view(np.float32))
    c = a.copy().view(np.uint32)
    c[1] = 0x7FC00000
    assert O.feature_key(a) == O.fnv1a(c.tobytes())


def test_quant_downf_matches_exact_fmod():
    rng = np.random.default_rng(0)
    x = (rng.random(20000) * 40).astype(np.float32)
    y = np.float32(0.0372)
    for xi in x[:2000]:
        q = O.quant_downf(float(xi), float(y))
        assert np.float32(q) == np.float32(xi - np.fmod(xi, y))


def test_closed_form_feature():
    # p1 = 0, n1 = x, p2 = x, n2 = y  ->  F = (1, 0, pi/2, pi/2) before quantization
    raw, q = O.compute_ppf([0, 0, 0], [1, 0, 0], [1, 0, 0], [0, 1, 0], 0.3)
    assert raw[0] == 1.0 and raw[1] == 0.0
    assert raw[2] == np.float32(math.pi / 2) and raw[3] == np.float32(math.pi / 2)
    # quant_downf(x, s) = fl(floor(x/s) * s): independent float64 evaluation
    s = np.float32(0.3)
    assert q[0] == np.float32(math.floor(1.0 / float(s)) * float(s))
    for k in (1, 2, 3):
        assert q[k] == np.float32(math.floor(float(raw[k]) / float(D)) * float(D))


def test_feature_rigid_invariance():
    rng = np.random.default_rng(1)
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(Q) < 0:
        Q[:, 0] *= -1
    t = rng.normal(size=3)
    for _ in range(200):
        p1, p2 = rng.normal(size=3), rng.normal(size=3)
        n1, n2 = rng.normal(size=3), rng.normal(size=3)
        n1 /= np.linalg.norm(n1)
        n2 /= np.linalg.norm(n2)
        r1, _ = O.compute_ppf(p1, n1, p2, n2, 0.1)
        r2, _ = O.compute_ppf(Q @ p1 + t, Q @ n1, Q @ p2 + t, Q @ n2, 0.1)
        assert np.allclose(r1, r2, atol=2e-5)


def _ulp_err(got, ref):
    got = got.astype(np.float64)
    r32 = ref.astype(np.float32)
    ulp = np.abs(np.spacing(r32).astype(np.float64))
    ulp[ulp == 0] = np.float64(np.finfo(np.float32).smallest_subnormal)
    return np.abs(got - ref) / ulp


def test_own_transcendentals_accuracy():
    rng = np.random.default_rng(2)
    x = np.linspace(-1, 1, 400001).astype(np.float32)
    assert _ulp_err(O.numerics(0, x), np.arccos(x.astype(np.float64))).max() <= 2.0
    y = (rng.normal(size=400000) * np.exp2(rng.integers(-10, 10, 400000))).astype(np.float32)
    xx = (rng.normal(size=400000) * np.exp2(rng.integers(-10, 10, 400000))).astype(np.float32)
    e = _ulp_err(O.numerics(1, y, xx), np.arctan2(y.astype(np.float64), xx.astype(np.float64)))
    assert e.max() <= 4.0
    t = np.linspace(-3.2, 3.2, 400001).astype(np.float32)
    assert _ulp_err(O.numerics(2, t), np.sin(t.astype(np.float64))).max() <= 2.0
    assert _ulp_err(O.numerics(3, t), np.cos(t.astype(np.float64))).max() <= 2.0


def test_atan2_special_values():
    pi = np.float32(3.141592654)
    cases = [((0.0, -1.0), pi), ((-0.0, -1.0), -pi), ((0.0, 1.0), 0.0), ((-0.0, 1.0), -0.0),
             ((0.0, -0.0), pi), ((-0.0, 0.0), -0.0), ((1.0, 0.0), np.float32(math.pi / 2)),
             ((-1.0, -0.0), -np.float32(math.pi / 2))]
    for (yy, xx), want in cases:
        got = O.numerics(1, np.array([yy], np.float32), np.array([xx], np.float32))[0]
        assert got == want and np.signbit(got) == np.signbit(np.float32(want)), (yy, xx, got)
    assert np.signbit(O.numerics(2, np.array([-0.0], np.float32))[0])


def test_alpha_bin_range_and_edges():
    rng = np.random.default_rng(3)
    a = rng.normal(size=(100000, 2)).astype(np.float32)
    b = rng.normal(size=(100000, 2)).astype(np.float32)
    bins = O.numerics(8, a.reshape(-1), b.reshape(-1))
    assert bins.min() >= 0 and bins.max() <= 30
    # compare with the float64 angle away from edges
    th = np.arctan2(a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0], (a * b).sum(1))
    pos = (th + float(np.float32(3.141592654))) / float(D)
    far = np.abs(pos - np.round(pos)) > 1e-3
    assert np.array_equal(bins[far], np.floor(pos[far]).astype(np.float32))
    # u antiparallel to v with cross = +0: alpha = +pi -> the single-point bin 30
    assert O.numerics(8, np.array([1, 0], np.float32), np.array([-1, 0], np.float32))[0] == 30
    # cross = -0 instead: alpha = -pi -> bin 0 (sign of zero decides, kernel.cu:338-340)
    assert O.numerics(8, np.array([1, -0.0], np.float32), np.array([-1, -0.0], np.float32))[0] == 0
