"""OMPL and vanilla state conventions (lib/osr/unit_world.cc:558-711), checked
through what the conventions mean for the robot's placement rather than by
restating the formulas:

  * a vanilla state places robot vertex v at R v + t, an OMPL state at
    R (v - O) + t with O the OMPL centre (unit_world.cc:591-604), so both
    must put every vertex at the same point;
  * an OMPL state and its unit-world translation place every robot vertex at
    points related by the scene's calibration matrix, also after
    enforceRobotCenter moved the robot's own centre away from O;
  * the alpha puzzle's OMPL solution path (res/alpha/alpha-1.2.org.path)
    stays collision free in the oracle after the translation, with and
    without a moved centre, and a translation with the centre correction's
    sign flipped does not;
  * unit -> OMPL -> unit and the angle-axis form are consistent.

CPU only: the translations are host arithmetic of the library."""
import os

import numpy as np
import pytest

from puzzletunneldiscovery_amd import pyosr

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(HERE, "golden", "ref")
ENV = os.path.join(REF, "alpha_env-1.2.org.obj")
ROB = os.path.join(REF, "alpha-1.2.org.obj")
MOVED = np.array([3.5, -7.25, 12.0])  # an enforced robot centre far from the mean


def rot(qw, qx, qy, qz):
    n = np.sqrt(qw * qw + qx * qx + qy * qy + qz * qz)
    w, x, y, z = qw / n, qx / n, qy / n, qz / n
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def ompl_states(n, seed, span=20.0):
    """Random OMPL states [t, qx, qy, qz, qw] with non-unit quaternions."""
    rng = np.random.default_rng(seed)
    q = np.zeros((n, 7))
    q[:, :3] = rng.uniform(-span, span, (n, 3))
    q[:, 3:] = rng.standard_normal((n, 4)) * rng.uniform(0.5, 2.0, (n, 1))
    return q


def alpha_world(center=None):
    uw = pyosr.UnitWorld()
    uw.loadModelFromFile(ENV)
    uw.loadRobotFromFile(ROB)
    if center is not None:
        uw.enforceRobotCenter(center)
    uw.scaleToUnit()
    uw.angleModel(0.0, 0.0)
    return uw


def robot_vertices(k=50, seed=0):
    V = []
    with open(ROB) as f:
        for line in f:
            if line.startswith("v "):
                V.append([float(x) for x in line.split()[1:4]])
    V = np.array(V)
    return V[np.random.default_rng(seed).choice(len(V), k, replace=False)]


def test_vanilla_and_ompl_place_vertices_alike():
    uw = alpha_world()
    O = uw.get_ompl_center(uw.GEO_ROB)
    V = robot_vertices()
    qv = ompl_states(200, 1)
    qo = uw.translate_vanilla_to_ompl(qv)
    assert np.array_equal(qo[:, 3:], qv[:, 3:])
    for a, b in zip(qv, qo):
        R = rot(a[6], a[3], a[4], a[5])  # w last
        pv = V @ R.T + a[:3]
        po = (V - O) @ R.T + b[:3]
        assert np.abs(pv - po).max() <= 1e-12 * (1 + np.abs(pv).max())
    back = uw.translate_ompl_to_vanilla(qo)
    assert np.abs(back - qv).max() <= 1e-12 * (1 + np.abs(qv).max())


@pytest.mark.parametrize("center", [None, MOVED])
def test_ompl_to_unit_preserves_placement(center):
    uw = alpha_world(center)
    O = uw.get_ompl_center(uw.GEO_ROB)
    S = uw.scene_matrix      # calibration: OMPL world -> unit world
    M = uw.robot_matrix      # robot vertices -> unit-scaled robot frame
    V = robot_vertices()
    Vh = np.hstack([V, np.ones((len(V), 1))])
    qo = ompl_states(200, 2)
    qu = uw.translate_ompl_to_unit(qo)
    qn = np.linalg.norm(qu[:, 3:], axis=1)
    assert np.abs(qn - 1).max() <= 4e-16
    for a, u in zip(qo, qu):
        R = rot(a[6], a[3], a[4], a[5])
        world = (V - O) @ R.T + a[:3]
        want = (np.hstack([world, np.ones((len(V), 1))]) @ S.T)[:, :3]
        Ru = rot(*u[3:])
        got = (Vh @ M.T)[:, :3] @ Ru.T + u[:3]
        # the calibration matrices are binary32, so the robot frame's offset
        # carries float rounding of centre * scale
        assert np.abs(got - want).max() <= 1e-6, np.abs(got - want).max()
    # unit -> OMPL inverts it
    back = uw.translate_unit_to_ompl(qu)
    qon = qo.copy()
    qon[:, 3:] /= np.linalg.norm(qo[:, 3:], axis=1, keepdims=True)
    sign = np.sign(np.sum(back[:, 3:] * qon[:, 3:], axis=1, keepdims=True))
    assert np.all(sign > 0)
    assert np.abs(back[:, :3] - qo[:, :3]).max() <= 1e-9 * 20
    assert np.abs(back[:, 3:] - qon[:, 3:]).max() <= 1e-15


def test_unit_to_ompl_angle_axis():
    uw = alpha_world(MOVED)
    qu = ompl_states(300, 3, span=1.0)
    qu[:, 3:] /= np.linalg.norm(qu[:, 3:], axis=1, keepdims=True)
    qu[:5, 4:] = 0.0            # identity rotations: angle 0, axis (1, 0, 0)
    qu[:5, 3] = 1.0
    quat = uw.translate_unit_to_ompl(qu)
    aa = uw.translate_unit_to_ompl(qu, to_angle_axis=True)
    assert np.array_equal(aa[:, :3], quat[:, :3])
    ang, ax = aa[:, 3], aa[:, 4:]
    assert np.all((ang >= 0) & (ang <= np.pi))
    assert np.array_equal(aa[:5, 3:], np.tile([0.0, 1.0, 0.0, 0.0], (5, 1)))
    assert np.abs(np.linalg.norm(ax, axis=1) - 1).max() <= 1e-15
    # the angle-axis pair rebuilds the quaternion up to its sign
    rebuilt = np.column_stack([np.sin(ang / 2)[:, None] * ax, np.cos(ang / 2)])  # w last
    s = np.sign(quat[:, 6:7])
    s[s == 0] = 1
    assert np.abs(rebuilt - s * quat[:, 3:]).max() <= 1e-14


def test_vanilla_to_unit_composes():
    uw = alpha_world(MOVED)
    qv = ompl_states(100, 4)
    assert np.array_equal(uw.translate_vanilla_to_unit(qv),
                          uw.translate_ompl_to_unit(uw.translate_vanilla_to_ompl(qv)))


def test_vanilla_points_to_unit():
    uw = alpha_world()
    S, M = uw.scene_matrix, uw.robot_matrix
    P = ompl_states(64, 5)[:, :3]
    Ph = np.hstack([P, np.ones((len(P), 1))])
    # environment points go through the calibration matrix (binary32 entries,
    # diagonal), which is exact here
    assert np.array_equal(uw.translate_vanilla_pts_to_unit(uw.GEO_ENV, P), (Ph @ S.T)[:, :3])
    # robot points relative to the OMPL centre: with the centre left at the
    # mean, that is the robot's own unit frame
    got = uw.translate_vanilla_pts_to_unit(uw.GEO_ROB, P)
    assert np.abs(got - (Ph @ M.T)[:, :3]).max() <= 1e-6
    assert uw.translate_vanilla_pts_to_unit(uw.GEO_ROB, np.zeros((0, 3))).shape == (0, 3)


def test_translations_need_a_robot():
    uw = pyosr.UnitWorld()
    uw.loadModelFromFile(ENV)
    with pytest.raises(RuntimeError):
        uw.translate_ompl_to_unit(np.zeros((1, 7)))


@pytest.mark.parametrize("center", [None, MOVED])
def test_alpha_solution_path_through_translate_ompl_to_unit(oracle_mod, center):
    """The OMPL solution path is free in the oracle world after the library's
    translation; the same translation with the centre correction's sign
    flipped puts the robot into the environment."""
    uw = alpha_world(center)
    ow = oracle_mod.OracleWorld()
    ow.load_obj(0, ENV)
    ow.load_obj(1, ROB)
    if center is not None:
        ow.enforce_robot_center(center)
    ow.scale_to_unit()
    ow.angle_model()
    path = np.loadtxt(os.path.join(REF, "alpha-1.2.org.path"))
    unit = uw.translate_ompl_to_unit(path)
    v = ow.valid_states(unit).astype(bool)
    assert v.all(), np.nonzero(~v)[0]
    if center is None:
        # no centre correction: the translation is translateToUnitState of the
        # reordered state
        w = path[:, [0, 1, 2, 6, 3, 4, 5]]
        direct = np.array([ow.translate_to_unit_state(q) for q in w])
        direct[:, 3:] /= np.linalg.norm(direct[:, 3:], axis=1, keepdims=True)
        assert np.abs(direct - unit).max() <= 1e-15
    else:
        d = (uw.get_ompl_center(uw.GEO_ROB).astype(np.float32) - np.float32(center).astype(np.float32)) \
            .astype(np.float64) * uw.scene_scale
        flipped = unit.copy()
        for i, u in enumerate(unit):
            flipped[i, :3] = u[:3] + 2 * rot(*u[3:]) @ d
        assert not ow.valid_states(flipped).astype(bool).all()


@pytest.mark.gpu
def test_gpu_alpha_path_with_moved_center(oracle_mod):
    """The engine agrees with the oracle on the translated path and on its
    sign-flipped copy, with the robot centre moved."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    uw = alpha_world(MOVED)
    ow = oracle_mod.OracleWorld()
    ow.load_obj(0, ENV)
    ow.load_obj(1, ROB)
    ow.enforce_robot_center(MOVED)
    ow.scale_to_unit()
    ow.angle_model()
    path = np.loadtxt(os.path.join(REF, "alpha-1.2.org.path"))
    unit = uw.translate_ompl_to_unit(path)
    assert uw.is_valid_states(unit).all()
    d = (uw.get_ompl_center(uw.GEO_ROB).astype(np.float32) - np.float32(MOVED).astype(np.float32)) \
        .astype(np.float64) * uw.scene_scale
    flipped = np.array([np.concatenate([u[:3] + 2 * rot(*u[3:]) @ d, u[3:]]) for u in unit])
    got = uw.is_valid_states(flipped)
    assert not got.all()
    assert np.array_equal(got, ow.valid_states(flipped).astype(bool))
