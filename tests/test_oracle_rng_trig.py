"""Pins for the oracle's RNG (random.glsl:10-31), fp32 sin/cos and sampling (random.glsl:33-62).

The reference has no tests; these checks fix the oracle against closed forms, hand-derived
known answers and mathematical properties (no check retypes the oracle's own formula).
"""
import math

import numpy as np
import pytest

import oracle as O

A = 747796405
M32 = 1 << 32


def test_seed_formula_known_answers():
    # (600*x + y) * (randomState + 1), u32 wrap (random.glsl:10)
    assert O.rng_seed(0, 0, 0) == 0
    assert O.rng_seed(1, 0, 0) == 600
    assert O.rng_seed(0, 600, 0) == 600          # pixel (0,600) aliases pixel (1,0): the literal 600
    assert O.rng_seed(3, 5, 1) == (1805 * 2)
    assert O.rng_seed(7158279, 0, 0) == (7158279 * 600) % M32  # wraps


def test_first_output_hand_derived():
    # state 0 -> step: 0*747796405 + 1 = 1
    # word = ((1 >> ((1 >> 28) + 4)) ^ 1) * 277803737 = 1 * 277803737 = 277803737 (0x108EF2D9)
    # word = (word >> 22) ^ word = 66 ^ 277803737 = 277803675 (0x108EF29B)
    # float(277803675) = 277803680.0 (fp32, 24-bit mantissa) ; / 2^32
    vals, st = O.rng_sequence(0, 1)
    assert st == 1
    assert vals[0] == np.float32(277803680.0 / 4294967296.0)


def test_lcg_has_full_period():
    """s -> a*s + 1 mod 2^32 (Hull-Dobell: c odd, a = 1 mod 4) returns to 0 after exactly 2^32 steps.
    Jump-ahead by repeated squaring of the affine map gives s_n without iterating."""
    assert A % 4 == 1

    def jump(n):
        acc_a, acc_c = 1, 0        # composed map x -> acc_a*x + acc_c
        cur_a, cur_c = A, 1
        while n:
            if n & 1:
                acc_a, acc_c = (cur_a * acc_a) % M32, (cur_a * acc_c + cur_c) % M32
            cur_a, cur_c = (cur_a * cur_a) % M32, (cur_a * cur_c + cur_c) % M32
            n >>= 1
        return acc_c  # image of 0
    assert jump(M32) == 0
    assert jump(M32 // 2) != 0
    # the oracle's step agrees with the jump-ahead for small n
    st = 0
    for n in range(1, 20):
        st = O.lib.orc_rng_step(st)
        assert st == jump(n)


def test_output_statistics():
    vals, _ = O.rng_sequence(12345, 200_000)
    assert vals.min() >= 0.0 and vals.max() <= 1.0
    assert abs(vals.mean() - 0.5) < 0.004
    hist, _ = np.histogram(vals, bins=64, range=(0.0, 1.0))
    chi2 = ((hist - len(vals) / 64) ** 2 / (len(vals) / 64)).sum()
    assert chi2 < 130  # 63 dof, p ~ 1e-6
    # consecutive pairs are not correlated
    assert abs(np.corrcoef(vals[:-1], vals[1:])[0, 1]) < 0.01


def test_float_to_uint_saturation():
    f = O.lib.orc_f2u_sat
    assert f(0.0) == 0 and f(-3.5) == 0 and f(float("nan")) == 0
    assert f(2.9) == 2
    assert f(4294967296.0) == 0xFFFFFFFF       # A2: alpha 1.0 * 4294967294.0f == 2^32 in fp32
    assert f(np.float32(0.5) * np.float32(4294967294.0)) == 2 ** 31


def _ulp_err(a, ref):
    a = np.float32(a)
    return abs(float(a) - ref) / float(np.spacing(np.float32(abs(ref)) if ref != 0 else np.float32(1e-30)))


def test_sin_cos_accuracy():
    xs = np.linspace(-2 * math.pi, 2 * math.pi, 20001).astype(np.float32)
    worst = 0.0
    for x in xs:
        s, c = O.lib.orc_sinf(float(x)), O.lib.orc_cosf(float(x))
        rs, rc = math.sin(float(x)), math.cos(float(x))
        # absolute error bound relative to 1 ulp at magnitude max(|ref|, 2^-8)
        for v, r in ((s, rs), (c, rc)):
            scale = float(np.spacing(np.float32(max(abs(r), 2.0 ** -8))))
            worst = max(worst, abs(v - r) / scale)
    assert worst <= 3.0
    assert O.lib.orc_sinf(0.0) == 0.0 and O.lib.orc_cosf(0.0) == 1.0
    for x in (0.3, 1.7, 4.0):
        assert O.lib.orc_sinf(-x) == -O.lib.orc_sinf(x)
        assert O.lib.orc_cosf(-x) == O.lib.orc_cosf(x)


def test_random_in_unit_sphere_closed_form():
    """(rho, theta, phi) uniform draws -> (rho sin phi cos theta, rho sin phi sin theta, rho cos phi),
    evaluated in float64 from the same draws (random.glsl:33-38)."""
    for seed in (1, 99, 123456789, 0xDEADBEEF):
        draws, _ = O.rng_sequence(seed, 3)
        v, st = O.random_in_unit_sphere(seed)
        rho = float(draws[0])
        theta = float(np.float32(2.0 * np.float32(np.pi)) * draws[1])
        phi = float(np.float32(np.pi) * draws[2])
        ref = np.array([rho * math.sin(phi) * math.cos(theta), rho * math.sin(phi) * math.sin(theta),
                        rho * math.cos(phi)])
        assert np.allclose(v, ref, atol=2e-6)
        assert np.linalg.norm(v) <= 1.0 + 1e-6
        _, st3 = O.rng_sequence(seed, 3)
        assert st == st3  # exactly three draws


def test_random_unit_vector_is_unit():
    norms = []
    for seed in range(1, 2000):
        v, _ = O.random_unit_vector(seed * 7919)
        if np.all(np.isfinite(v)):
            norms.append(np.linalg.norm(v.astype(np.float64)))
    assert np.allclose(norms, 1.0, atol=1e-6)


def test_reflect_refract_schlick():
    rng = np.random.default_rng(3)
    for _ in range(200):
        i = rng.normal(size=3).astype(np.float32)
        i /= np.linalg.norm(i)
        n = rng.normal(size=3).astype(np.float32)
        n /= np.linalg.norm(n)
        if np.dot(i, n) > 0:
            n = -n
        r = np.zeros(3, np.float32)
        O.lib.orc_reflect(i.ctypes.data, n.ctypes.data, r.ctypes.data)
        assert abs(np.linalg.norm(r) - 1.0) < 1e-5
        assert abs(np.dot(r, n) + np.dot(i, n)) < 1e-5
        eta = float(rng.uniform(0.5, 1.5))
        t = np.zeros(3, np.float32)
        O.lib.orc_refract(i.ctypes.data, n.ctypes.data, eta, t.ctypes.data)
        sin_i = np.linalg.norm(np.cross(i, n))
        if eta * sin_i > 1.0 + 1e-5:
            assert np.all(t == 0)  # total internal reflection
        elif eta * sin_i < 1.0 - 1e-3:
            assert abs(np.linalg.norm(np.cross(t, n)) - eta * sin_i) < 1e-4  # Snell
            assert np.dot(t, n) < 0
    for ri in (0.5, 1.0 / 1.5, 1.5):
        r0 = ((1 - ri) / (1 + ri)) ** 2
        assert abs(O.lib.orc_schlick(1.0, ri) - r0) < 1e-7
        assert abs(O.lib.orc_schlick(0.0, ri) - 1.0) < 1e-6
