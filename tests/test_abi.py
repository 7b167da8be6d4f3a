"""The C ABI (include/se3cq.h) without a GPU: the library loads, exports every
declared entry point, and its host-side work -- mesh ingest, unit-world
normalisation (lib/osr/unit_world.cc:94-161), state arithmetic
(lib/osr/osr_state.cc) and argument checking -- matches the oracle bit for
bit.  Nothing here launches a kernel."""
import ctypes
import os
import re

import numpy as np
import pytest

from puzzletunneldiscovery_amd import _native as N
from puzzletunneldiscovery_amd import pyosr
from puzzletunneldiscovery_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "tests", "golden", "ref")


def _declared(header="se3cq.h", prefix="se3cq_"):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(" + prefix + r"[a-z0-9_]+)\s*\(", text)))


def _gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def test_every_declared_symbol_is_exported():
    L = N.lib()
    names = _declared()
    assert len(names) >= 40
    for name in names:
        assert hasattr(L, name), name
        assert name in N.SIGNATURES, name
    assert set(N.SIGNATURES) == set(names)
    assert N.lib().se3cq_version().decode()
    mlp = _declared("se3mlp.h", "se3mlp_")
    assert len(mlp) >= 9 and set(N.MLP_SIGNATURES) == set(mlp)
    for name in mlp:
        assert hasattr(L, name), name


def _pair(oracle_mod, env, rob, obj=False):
    ow = oracle_mod.OracleWorld()
    uw = pyosr.UnitWorld()
    if obj:
        ow.load_obj(0, env)
        ow.load_obj(1, rob)
        uw.loadModelFromFile(env)
        uw.loadRobotFromFile(rob)
    else:
        ow.load_arrays(0, *env)
        ow.load_arrays(1, *rob)
        uw.load_mesh(0, *env)
        uw.load_mesh(1, *rob)
    return ow, uw


def _normalise(ow, uw):
    ow.scale_to_unit()
    ow.angle_model()
    uw.scaleToUnit()
    uw.angleModel(0.0, 0.0)


@pytest.mark.parametrize("case", ["torus", "alpha_obj", "alpha2_obj"])
def test_host_geometry_matches_oracle(oracle_mod, torus_meshes, case):
    if case == "torus":
        ow, uw = _pair(oracle_mod, *torus_meshes)
    elif case == "alpha_obj":
        ow, uw = _pair(oracle_mod, os.path.join(REF, "alpha_env-1.2.org.obj"),
                       os.path.join(REF, "alpha-1.2.org.obj"), obj=True)
    else:
        ow, uw = _pair(oracle_mod, os.path.join(REF, "env-1.2.obj"),
                       os.path.join(REF, "robot.obj"), obj=True)
    for geo in (0, 1):
        Vo, Fo = ow.mesh(geo)
        Vu, Fu = uw.mesh(geo)
        assert np.array_equal(Vo, Vu) and np.array_equal(Fo, Fu)
        assert np.array_equal(ow.center(geo).astype(np.float32), uw.get_ompl_center(geo).astype(np.float32))
    _normalise(ow, uw)
    assert np.float32(ow.scene_scale) == np.float32(uw.scene_scale)
    for geo in (0, 1):
        assert np.array_equal(ow.cd_vertices(geo), uw.cd_vertices(geo))
    assert np.array_equal(ow.calibration(), uw.scene_matrix)
    s = np.float32(uw.scene_scale)
    c = uw.get_ompl_center(1).astype(np.float32)
    M = uw.robot_matrix   # scale(s) * translate(-centre), binary32 (unit_world.cc:152-154)
    assert np.array_equal(M[:3, :3], np.eye(3) * s)
    assert np.array_equal(M[:3, 3], (s * -c).astype(np.float64)) and np.array_equal(M[3], [0, 0, 0, 1])
    qs = W.random_states(300, seed=5, lo=-2.0, hi=3.0)
    want = np.array([ow.translate_to_unit_state(q) for q in qs])
    assert np.array_equal(want, np.array([uw.translate_to_unit_state(q) for q in qs]))


def test_enforced_center(oracle_mod, torus_meshes):
    ow, uw = _pair(oracle_mod, *torus_meshes)
    c = [0.01, -0.02, 0.03]
    ow.enforce_robot_center(c)
    uw.enforceRobotCenter(c)
    _normalise(ow, uw)
    assert np.array_equal(ow.cd_vertices(1), uw.cd_vertices(1))


def test_state_arithmetic_matches_oracle(oracle_mod):
    """distance / interpolate / apply / decompose (osr_state.cc:27-44,94-99,
    130-154,243-259) use the same portable transcendental functions in the
    library and the oracle, so they agree to the bit."""
    rng = np.random.default_rng(9)
    a = W.random_states(400, seed=6)
    b = W.random_states(400, seed=7)
    b[:50] = a[:50]                              # distance 0
    b[50:100, 3:] = -a[50:100, 3:]               # antipodal quaternions
    b[100:150, 3:] = a[100:150, 3:] + 1e-9       # nearly parallel (lerp branch)
    for i in range(400):
        assert pyosr.distance(a[i], b[i]) == oracle_mod.distance(a[i], b[i])
        tau = float(rng.uniform(-0.2, 1.2))
        assert np.array_equal(pyosr.interpolate(a[i], b[i], tau), oracle_mod.interpolate(a[i], b[i], tau))
    md = pyosr.multi_distance(a[0], b)
    assert np.array_equal(md, [oracle_mod.distance(a[0], x) for x in b])
    for i in range(100):
        tr = rng.normal(size=3) * 0.1
        aa = rng.normal(size=3) * rng.uniform(0, 3)
        assert np.array_equal(pyosr.apply(a[i], tr, aa), oracle_mod.apply(a[i], tr, aa))
        # decompose_2 normalises first (Eigen: q / sqrt((x^2 + z^2) + (y^2 + w^2)))
        t, R = pyosr.decompose_2(a[i])
        w, x, y, z = a[i, 3:]
        qn = a[i].copy()
        qn[3:] = a[i, 3:] / np.sqrt((x * x + z * z) + (y * y + w * w))
        Ro, to = oracle_mod.state_to_transform(qn)
        assert np.array_equal(R, Ro) and np.array_equal(t, to)
    # apply over rows, with touch-sampling sized moves (unit translation, up to 2 pi)
    trs = rng.normal(size=(300, 3))
    trs /= np.linalg.norm(trs, axis=1, keepdims=True)
    aas = rng.uniform(-1, 1, (300, 3)) * 2 * np.pi
    aas[:5] = 0.0
    got = pyosr.apply_batch(np.vstack([a, a])[:300], trs, aas)
    assert np.array_equal(got, [oracle_mod.apply(q, t, x) for q, t, x in zip(np.vstack([a, a])[:300], trs, aas)])
    assert pyosr.apply_batch(np.zeros((0, 7)), np.zeros((0, 3)), np.zeros((0, 3))).shape == (0, 7)
    with pytest.raises(ValueError):
        pyosr.apply_batch(a[:3], trs[:2], aas[:3])


def test_compose_from_angleaxis_closed_form():
    """Eigen's AngleAxis -> quaternion: (cos(a/2), sin(a/2) * axis)."""
    q = pyosr.compose_from_angleaxis([1, 2, 3], np.pi / 2, [0, 0, 1])
    assert np.allclose(q, [1, 2, 3, np.cos(np.pi / 4), 0, 0, np.sin(np.pi / 4)], atol=1e-15)
    assert np.array_equal(pyosr.get_identity_state_vector(), [0, 0, 0, 1, 0, 0, 0])


def test_decompose_3_closed_forms():
    """decompose_3 (osr_state.cc:101-107): the state built from (angle, axis)
    with angle in [0, pi] gives them back (within a few ulp: the angle goes
    through atan2 and a square root); a quaternion with w < 0 is the same
    rotation with axis and vector part negated; the identity and a rotation by
    pi keep Eigen's conventions; the translation passes unchanged and the
    quaternion is normalised first."""
    rng = np.random.default_rng(12)
    for _ in range(300):
        ang = float(rng.uniform(1e-3, np.pi - 1e-3))
        ax = rng.normal(size=3)
        ax /= np.linalg.norm(ax)
        tr = rng.normal(size=3)
        q = np.concatenate([tr, [np.cos(ang / 2)], np.sin(ang / 2) * ax])
        t, a, u = pyosr.decompose_3(q)
        assert np.array_equal(t, tr)
        assert abs(a - ang) <= 8 * np.spacing(np.pi) and np.allclose(u, ax, atol=1e-14)
        qn = q.copy()
        qn[3:] *= -float(rng.uniform(0.5, 2.0))      # same rotation, w < 0, not unit
        t2, a2, u2 = pyosr.decompose_3(qn)
        assert abs(a2 - ang) <= 8 * np.spacing(np.pi) and np.allclose(u2, ax, atol=1e-14)
        assert np.linalg.norm(u2) == pytest.approx(1.0, abs=4e-16)
    t, a, u = pyosr.decompose_3([1, 2, 3, 1, 0, 0, 0])
    assert a == 0.0 and np.array_equal(u, [1.0, 0.0, 0.0])
    t, a, u = pyosr.decompose_3([0, 0, 0, -2.0, 0, 0, 0])
    assert a == 0.0 and np.array_equal(u, [1.0, 0.0, 0.0])
    t, a, u = pyosr.decompose_3([0, 0, 0, 0.0, 0.0, 0.0, 1.0])   # rotation by pi about z
    assert a == 2 * pyosr.distance([0] * 3 + [1, 0, 0, 0], [0] * 3 + [0, 0, 0, 1]) / 2
    assert abs(a - np.pi) <= np.spacing(np.pi) and np.array_equal(u, [0.0, 0.0, 1.0])
    # a tiny rotation: the axis still comes out of the vector part, scaled
    t, a, u = pyosr.decompose_3([0, 0, 0, 1.0, 3e-17, 4e-17, 0.0])
    assert a == pytest.approx(1e-16, rel=1e-12) and np.allclose(u, [0.6, 0.8, 0.0], atol=1e-15)


def test_errors_are_reported_not_thrown():
    L = N.lib()
    h = ctypes.c_void_p()
    assert L.se3cq_world_create(ctypes.byref(h)) == 0
    try:
        assert L.se3cq_load_file(h, 0, b"/nonexistent/x.obj") == -3
        assert b"/nonexistent/x.obj" in L.se3cq_last_error()
        V = np.zeros((3, 3), np.float32)
        F = np.array([[0, 1, 5]], np.uint32)
        assert L.se3cq_load_mesh(h, 0, V.ctypes.data_as(N.c_float_p), 3, F.ctypes.data_as(N.c_u32_p), 1) == -1
        assert L.se3cq_load_mesh(h, 7, V.ctypes.data_as(N.c_float_p), 3, F.ctypes.data_as(N.c_u32_p), 1) == -1
        V[0, 0] = np.inf
        F[0, 2] = 2
        assert L.se3cq_load_mesh(h, 0, V.ctypes.data_as(N.c_float_p), 3, F.ctypes.data_as(N.c_u32_p), 1) == -1
        assert b"vertex" in L.se3cq_last_error()
        assert L.se3cq_angle_model(h, ctypes.c_float(0.5), ctypes.c_float(0.0)) != 0
    finally:
        assert L.se3cq_world_destroy(h) == 0
    w = pyosr.UnitWorld()
    with pytest.raises(ValueError):
        w.transit_states_to_with_contact(np.zeros((3, 7)), np.zeros((2, 7)), 0.01)
    with pytest.raises(RuntimeError):
        w.loadModelFromFile("/nonexistent/y.obj")


def test_unported_members_raise():
    w = pyosr.UnitWorld()
    with pytest.raises(NotImplementedError):
        w.transit_state(pyosr.get_identity_state_vector(), 0, 0.1, 0.01)
    # isDisentangled without models throws in the reference (unit_world.cc:251-253)
    with pytest.raises(RuntimeError, match="not loaded"):
        w.is_disentangled(pyosr.get_identity_state_vector())


@pytest.mark.skipif(_gpu(), reason="checks the no-device error path")
def test_queries_fail_loudly_without_a_device(torus_meshes):
    """No CPU fallback: without a HIP device a query raises."""
    w = pyosr.create_unit_world_from_arrays(*torus_meshes)
    with pytest.raises(RuntimeError, match="HIP device"):
        w.is_valid_states(W.random_states(8, seed=1))


@pytest.mark.parametrize("vd", [0.0, -1.0 / 1024, float("nan"), float("inf")])
def test_verify_step_must_be_positive(torus_meshes, vd):
    """A verify step that is not positive and finite never advances the
    reference's stepping loop (unit_world.cc:448-470); every segment entry
    point rejects it before any device work."""
    w = pyosr.create_unit_world_from_arrays(*torus_meshes)
    a = W.random_states(4, seed=2)
    b = W.random_states(4, seed=3)
    calls = [
        lambda: w.transit_states_to_with_contact(a, b, vd),
        lambda: w.is_valid_transitions(a, b, vd),
        lambda: w.touch_search(a, b, vd, 20),
        lambda: w.calculate_visibility_matrix(a, True, vd),
        lambda: w.calculate_visibility_matrix2(a, True, b, True, vd),
        lambda: w.calculate_visibility_pair(a, True, b, True, vd),
        lambda: w.batch(a, b, vd, a),
    ]
    for call in calls:
        with pytest.raises(RuntimeError, match="verify step"):
            call()


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus above 1 exits with a message, before
    it touches torch or a device."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and "--gpus 2" in r.stderr
