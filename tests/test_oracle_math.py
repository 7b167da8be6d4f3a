"""Pins the portable transcendental functions and the oracle's state
arithmetic (lib/osr/osr_state.cc) against mathematics: mpmath references,
closed forms and invariants."""
import math

import mpmath as mp
import numpy as np
import pytest

mp.mp.prec = 120


def cr(f, *args):
    """The correctly rounded (nearest, ties to even) binary64 value of f.
    mpmath has no signed zero, so zero arguments take math's value."""
    if all(float(a) == 0.0 for a in args):
        return {mp.sin: math.sin, mp.cos: math.cos, mp.acos: math.acos, mp.atan2: math.atan2}[f](*args)
    with mp.workprec(320):
        return float(f(*[mp.mpf(float(a)) for a in args]))


def same(a, b):
    """Bitwise equality of two binary64 values (NaN equals NaN)."""
    return np.float64(a).tobytes() == np.float64(b).tobytes() or (a != a and b != b)


def _sin_args(rng):
    k = np.arange(1, 3001)
    near = np.nextafter(k * (math.pi / 2), np.inf) + rng.integers(-3, 4, len(k)) * np.spacing(k * (math.pi / 2))
    xs = np.concatenate([
        rng.uniform(-math.pi, math.pi, 3000),
        rng.uniform(0.0, math.pi / 2, 2000),                                   # slerp weights
        np.ldexp(rng.uniform(1, 2, 2000), rng.integers(-30, 20, 2000)) * rng.choice([-1, 1], 2000),
        near,                                                                  # next to multiples of pi/2
        np.ldexp(rng.uniform(1, 2, 200), rng.integers(-40, -27, 200)),        # tiny
        [0.0, -0.0, math.pi / 4, -math.pi / 4, math.pi / 2, 2.0 ** -27, 2.0 ** 20 * 0.999999],
    ])
    return xs


def test_sin_cos_correctly_rounded(oracle_mod):
    """se3_sin and se3_cos round correctly: bit-equal to mpmath's value
    rounded to nearest, on random arguments, the slerp domain [0, pi/2],
    magnitudes 2^-30 .. 2^20, arguments next to multiples of pi/2 (where the
    reduction cancels most bits) and tiny ones."""
    L = oracle_mod.lib()
    for x in _sin_args(np.random.default_rng(1)):
        x = float(x)
        assert same(L.or_se3_sin(x), cr(mp.sin, x)), ("sin", x)
        assert same(L.or_se3_cos(x), cr(mp.cos, x)), ("cos", x)


def test_acos_correctly_rounded(oracle_mod):
    L = oracle_mod.lib()
    rng = np.random.default_rng(2)
    xs = np.concatenate([
        rng.uniform(-1, 1, 4000),
        1 - np.ldexp(rng.uniform(0, 1, 1000), -rng.integers(1, 53, 1000)),    # next to 1 (slerp's |dot|)
        -1 + np.ldexp(rng.uniform(0, 1, 300), -rng.integers(1, 53, 300)),
        np.ldexp(rng.uniform(-1, 1, 300), -rng.integers(0, 80, 300)),         # next to 0
        [0.0, -0.0, 0.5, -0.5, 1.0, -1.0, np.nextafter(1.0, 0.0), np.nextafter(-1.0, 0.0)],
    ])
    for x in xs:
        x = float(x)
        assert same(L.or_se3_acos(x), cr(mp.acos, x)), x
    assert math.isnan(L.or_se3_acos(1.5)) and math.isnan(L.or_se3_acos(float("nan")))


def test_atan2_correctly_rounded(oracle_mod):
    """All four quadrants, ratios from 2^-80 to 2^80 (angularDistance calls
    atan2(|vec|, |w|) with both in [0, 1]), and C99's special values."""
    L = oracle_mod.lib()
    rng = np.random.default_rng(3)
    ys = np.concatenate([rng.uniform(-2, 2, 3000), rng.uniform(0, 1, 2000),
                         np.ldexp(rng.uniform(1, 2, 2000), rng.integers(-80, 80, 2000)) * rng.choice([-1, 1], 2000)])
    xs = np.concatenate([rng.uniform(-2, 2, 3000), rng.uniform(0, 1, 2000),
                         rng.uniform(-2, 2, 2000)])
    for y, x in zip(ys, xs):
        y, x = float(y), float(x)
        assert same(L.or_se3_atan2(y, x), cr(mp.atan2, y, x)), (y, x)
    inf, nan = float("inf"), float("nan")
    for y in (0.0, -0.0, 1.5, -1.5, inf, -inf, nan, 1e-310, -1e300):
        for x in (0.0, -0.0, 2.0, -2.0, inf, -inf, nan, 1e-310, -1e300):
            assert same(L.or_se3_atan2(y, x), math.atan2(y, x)), (y, x)


def test_slow_paths_correctly_rounded(oracle_mod):
    """The 256-bit fixed-point evaluations that decide when the double-double
    one cannot (about once in 2^40 calls) are correctly rounded themselves;
    they are called here directly."""
    L = oracle_mod.lib()
    rng = np.random.default_rng(4)
    for x in np.concatenate([rng.uniform(-4, 4, 400), np.ldexp(rng.uniform(1, 2, 200), rng.integers(-26, 19, 200))]):
        x = float(x)
        if abs(x) < 2.0 ** -26:
            continue
        assert same(L.or_cr_sin_slow(x), cr(mp.sin, x)), x
        assert same(L.or_cr_cos_slow(x), cr(mp.cos, x)), x
    for x in np.concatenate([rng.uniform(-0.999, 0.999, 400), 1 - np.ldexp(rng.uniform(0, 1, 100), -rng.integers(1, 50, 100))]):
        x = float(x)
        if x != 0.0:
            assert same(L.or_cr_acos_slow(x), cr(mp.acos, x)), x
    for y, x in zip(np.ldexp(rng.uniform(1, 2, 400), rng.integers(-29, 3, 400)) * rng.choice([-1, 1], 400),
                    rng.uniform(-3, 3, 400)):
        y, x = float(y), float(x)
        r = min(abs(y), abs(x)) / max(abs(y), abs(x))
        if r >= 2.0 ** -30:
            assert same(L.or_cr_atan2_slow(y, x), cr(mp.atan2, y, x)), (y, x)


def test_host_libm_mode_differs_only_in_last_bit(oracle_mod):
    """Libm mode 1 calls the host's sin/cos/acos/atan2 (glibc's are within an
    ulp and usually correctly rounded): each value is within one ulp of the
    correctly rounded one and differs from it rarely."""
    L = oracle_mod.lib()
    rng = np.random.default_rng(5)
    xs = rng.uniform(-math.pi, math.pi, 20000)
    diff = 0
    for x in xs:
        x = float(x)
        for mine, host in ((L.or_se3_sin(x), math.sin(x)), (L.or_se3_cos(x), math.cos(x))):
            assert abs(mine - host) <= np.spacing(abs(mine)), x
            diff += mine != host
    assert diff < 0.01 * 2 * len(xs)


def _rand_states(n, seed):
    from puzzletunneldiscovery_amd import workloads as W
    return W.random_states(n, seed=seed)


def test_rotation_closed_form(oracle_mod):
    """toRotationMatrix of q = n*u (u unit) equals n^2 R(u) + (1 - n^2) I, and
    R(u) is orthonormal with determinant 1."""
    rng = np.random.default_rng(4)
    for q in _rand_states(200, 5):
        R, t = oracle_mod.state_to_transform(q)
        assert np.array_equal(t, q[:3])
        u = q[3:] / np.linalg.norm(q[3:])
        Ru, _ = oracle_mod.state_to_transform(np.concatenate([q[:3], u]))
        assert np.allclose(Ru.T @ Ru, np.eye(3), atol=1e-14)
        assert abs(np.linalg.det(Ru) - 1.0) < 1e-14
        n = rng.uniform(0.5, 2.0)
        Rn, _ = oracle_mod.state_to_transform(np.concatenate([q[:3], n * u]))
        assert np.allclose(Rn, n * n * Ru + (1 - n * n) * np.eye(3), atol=1e-13)
        # textbook formula with w first
        w, x, y, z = u
        ref = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
        assert np.allclose(Ru, ref, atol=1e-15)


def test_distance_closed_form(oracle_mod):
    qs = _rand_states(300, 6)
    for a, b in zip(qs[:-1], qs[1:]):
        d = oracle_mod.distance(a, b)
        ua, ub = a[3:] / np.linalg.norm(a[3:]), b[3:] / np.linalg.norm(b[3:])
        ang = 2 * math.acos(min(1.0, abs(float(np.dot(ua, ub)))))
        assert abs(d - (np.linalg.norm(a[:3] - b[:3]) + ang)) < 1e-7
        assert abs(d - oracle_mod.distance(b, a)) < 1e-14
        assert oracle_mod.distance(a, a) == 0.0


def test_interpolate_endpoints_and_slerp(oracle_mod):
    qs = _rand_states(300, 7)
    for a, b in zip(qs[:-1], qs[1:]):
        ua, ub = a[3:] / np.linalg.norm(a[3:]), b[3:] / np.linalg.norm(b[3:])
        s0 = oracle_mod.interpolate(a, b, 0.0)
        assert np.array_equal(s0[:3], a[:3])
        assert np.allclose(s0[3:], ua, atol=1e-16)
        s1 = oracle_mod.interpolate(a, b, 1.0)
        assert np.array_equal(s1[:3], b[:3])
        sign = 1.0 if np.dot(ua, ub) >= 0 else -1.0
        assert np.allclose(s1[3:], sign * ub, atol=1e-15)
        # slerp: unit norm, constant angular speed along the shorter arc
        theta = math.acos(min(1.0, abs(float(np.dot(ua, ub)))))
        for tau in (0.25, 0.5, 0.8):
            s = oracle_mod.interpolate(a, b, tau)
            assert abs(np.linalg.norm(s[3:]) - 1) < 1e-14
            assert abs(math.acos(min(1.0, abs(float(np.dot(s[3:], ua))))) - tau * theta) < 1e-7
            assert np.allclose(s[:3], a[:3] * (1 - tau) + b[:3] * tau, atol=1e-15)


def test_apply_closed_form(oracle_mod):
    qs = _rand_states(100, 8)
    rng = np.random.default_rng(9)
    for q in qs:
        tr = rng.normal(size=3)
        aa = rng.normal(size=3)
        out = oracle_mod.apply(q, tr, aa)
        assert np.allclose(out[:3], q[:3] + tr, atol=1e-15)
        ang = np.linalg.norm(aa)
        ax = aa / ang
        qa = np.concatenate([[math.cos(ang / 2)], math.sin(ang / 2) * ax])
        u = q[3:] / np.linalg.norm(q[3:])
        w1, x1, y1, z1 = qa
        w2, x2, y2, z2 = u
        ref = np.array([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                        w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2])
        assert np.allclose(out[3:], ref, atol=1e-14)
    out = oracle_mod.apply(qs[0], np.zeros(3), np.zeros(3))
    assert np.allclose(out[3:], qs[0][3:] / np.linalg.norm(qs[0][3:]), atol=1e-15)
