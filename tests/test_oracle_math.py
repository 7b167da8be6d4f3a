"""Pins the portable transcendental functions and the oracle's state
arithmetic (lib/osr/osr_state.cc) against mathematics: mpmath references,
closed forms and invariants."""
import math

import mpmath as mp
import numpy as np
import pytest

mp.mp.prec = 120


def ulps(got, want):
    """Distance in units of the last place of the correctly rounded value."""
    want_d = float(want)
    if want_d == 0.0:
        return abs(got) / 5e-324
    u = math.ulp(want_d)
    return abs(mp.mpf(got) - want) / u


@pytest.mark.parametrize("fn,ref,lo,hi", [
    ("or_se3_sin", mp.sin, -12.0, 12.0),
    ("or_se3_cos", mp.cos, -12.0, 12.0),
    ("or_se3_acos", mp.acos, -1.0, 1.0),
])
def test_unary_accuracy(oracle_mod, fn, ref, lo, hi):
    f = getattr(oracle_mod.lib(), fn)
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(lo, hi, 3000), rng.uniform(-1e-3, 1e-3, 300),
                         [lo, hi, 0.0, 0.5, -0.5, math.pi / 4, -math.pi / 4, math.pi / 2]])
    xs = xs[(xs >= lo) & (xs <= hi)]
    worst = max(ulps(f(float(x)), ref(mp.mpf(float(x)))) for x in xs if abs(float(ref(mp.mpf(float(x))))) > 1e-300)
    assert worst <= 2.0, (fn, worst)


def test_atan2_accuracy(oracle_mod):
    f = oracle_mod.lib().or_se3_atan2
    rng = np.random.default_rng(2)
    ys = rng.uniform(-2, 2, 3000)
    xs = rng.uniform(-2, 2, 3000)
    worst = 0.0
    for y, x in zip(ys, xs):
        worst = max(worst, ulps(f(float(y), float(x)), mp.atan2(mp.mpf(float(y)), mp.mpf(float(x)))))
    # the (z - 1) / (z + 1) reduction adds up to ~1 ulp above tan(pi/8)
    assert worst <= 3.0
    assert f(0.0, 0.0) == 0.0 and f(0.0, 1.0) == 0.0
    assert abs(f(1.0, 0.0) - math.pi / 2) <= math.ulp(math.pi / 2)


def test_against_glibc(oracle_mod):
    """The portable functions and glibc agree to within 2 ulps (they are
    different implementations of the same functions)."""
    L = oracle_mod.lib()
    rng = np.random.default_rng(3)
    xs = rng.uniform(0, math.pi / 2, 5000)
    for fn, g in (("or_se3_sin", math.sin), ("or_se3_cos", math.cos)):
        f = getattr(L, fn)
        d = [abs(f(float(x)) - g(float(x))) / math.ulp(g(float(x))) for x in xs]
        assert max(d) <= 2.0
    xs = rng.uniform(0, 1, 5000)
    d = [abs(L.or_se3_acos(float(x)) - math.acos(float(x))) / math.ulp(math.acos(float(x))) for x in xs]
    assert max(d) <= 2.0


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
