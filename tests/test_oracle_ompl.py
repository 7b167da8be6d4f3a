"""The oracle's restatement of OMPL's SE(3) space and discrete motion
validator (OMPL 1.4.1, which the reference's ompl.app fork builds on; OMPL
itself is not in the reference's tree), pinned by closed forms and by the
reference's own OMPL solution path, plus the host-side rules of the
OmplDriver mirror (lib/pyse3ompl/ompldriver.h:70-120).  CPU only."""
import math
import os

import numpy as np
import pytest

from puzzletunneldiscovery_amd import pyosr, pyse3ompl

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(HERE, "golden", "ref")
ENV = os.path.join(REF, "alpha_env-1.2.org.obj")
ROB = os.path.join(REF, "alpha-1.2.org.obj")
MOVED = np.array([3.5, -7.25, 12.0])
# the environment's and the solution path's bounding boxes, widened by 10
LO = np.array([-11.892796, -13.170522, -15.239033])
HI = np.array([68.73492, 75.37971186, 56.18796844])


def quat_axis_angle(axis, angle):
    """OMPL order (x, y, z, w)."""
    a = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    return np.concatenate([math.sin(angle / 2) * a, [math.cos(angle / 2)]])


def state(t, q):
    return np.concatenate([np.asarray(t, np.float64), q])


@pytest.fixture(scope="module")
def alpha(oracle_mod):
    ow = oracle_mod.OracleWorld()
    ow.load_obj(0, ENV)
    ow.load_obj(1, ROB)
    ow.scale_to_unit()
    ow.angle_model()
    return ow


@pytest.fixture(scope="module")
def path():
    return np.loadtxt(os.path.join(REF, "alpha-1.2.org.path"))


def test_longest_segment_closed_form(oracle_mod):
    O = oracle_mod
    lo, hi = np.zeros(3), np.array([3.0, 4.0, 12.0])  # diagonal 13
    for res in (0.01, 1e-4, 0.5):
        want = (13.0 + math.pi / 2) * res
        assert O.ompl_longest_segment(lo, hi, res) == pytest.approx(want, rel=1e-15)
        assert pyse3ompl.longest_valid_segment(lo, hi, res) == O.ompl_longest_segment(lo, hi, res)
    for bad in (0.0, 1.0, -0.1, math.nan, 2.0):
        with pytest.raises(RuntimeError):
            pyse3ompl.longest_valid_segment(lo, hi, bad)
    with pytest.raises(RuntimeError):
        pyse3ompl.longest_valid_segment(hi, lo, 0.01)


def test_distance_closed_forms(oracle_mod):
    O = oracle_mod
    I = quat_axis_angle([1, 0, 0], 0.0)
    assert O.ompl_distance(state([0, 0, 0], I), state([3, 4, 0], I)) == 5.0
    for ang in (0.1, 1.0, 2.5, 3.1):
        q = quat_axis_angle([0.3, -1.0, 0.5], ang)
        # SO(3) arc length of OMPL is the quaternion angle, half the rotation
        assert O.ompl_distance(state([1, 2, 3], I), state([1, 2, 3], q)) == pytest.approx(ang / 2, abs=1e-12)
        # q and -q are one rotation
        assert O.ompl_distance(state([0, 0, 0], q), state([0, 0, 0], -q)) == 0.0
    # |dot| within 1e-9 of 1 counts as no rotation (MAX_QUATERNION_NORM_ERROR)
    tiny = quat_axis_angle([0, 0, 1], 4e-5)
    assert O.ompl_distance(state([0, 0, 0], I), state([0, 0, 0], tiny)) == 0.0
    assert O.ompl_distance(state([0, 0, 0], I), state([0, 0, 0], quat_axis_angle([0, 0, 1], 1e-4))) > 0


def test_interpolation_closed_forms(oracle_mod):
    O = oracle_mod
    a = state([1.0, -2.0, 0.5], quat_axis_angle([0, 0, 1], 0.0))
    for ang in (0.3, 2.0, 3.0):
        b = state([4.0, 2.0, -1.5], quat_axis_angle([0, 0, 1], ang))
        for t in (0.0, 0.25, 0.5, 0.9, 1.0):
            s = O.ompl_interpolate(a, b, t)
            assert np.array_equal(s[:3], a[:3] + (b[:3] - a[:3]) * t)
            # constant angular velocity about the common axis
            assert np.abs(s[3:] - quat_axis_angle([0, 0, 1], t * ang)).max() <= 2e-15
        # the long-arc case: -q2 is the same rotation, and the slerp takes
        # the short way to it
        bn = b.copy()
        bn[3:] = -bn[3:]
        for t in (0.25, 0.5, 0.75):
            s1, s2 = O.ompl_interpolate(a, b, t), O.ompl_interpolate(a, bn, t)
            assert np.abs(s1[3:] - s2[3:]).max() <= 2e-15
    # below the arc-length threshold the rotation stays at `from`
    b = state([1, 1, 1], quat_axis_angle([0, 1, 0], 3e-5))
    assert np.array_equal(O.ompl_interpolate(a, b, 0.5)[3:], a[3:])


def random_ompl_states(n, rng, lo=LO + 10, hi=HI - 10):
    q = np.zeros((n, 7))
    q[:, :3] = rng.uniform(lo, hi, (n, 3))
    r = rng.standard_normal((n, 4))
    q[:, 3:] = r / np.linalg.norm(r, axis=1, keepdims=True)
    return q


def test_check_motion_tests_exactly_the_validator_samples(alpha, oracle_mod):
    """checkMotion (midpoint-first interval queue) decides as 'the end state
    and every interpolate(k / nd), 0 < k < nd, are free', with nd =
    ceil(distance / longest segment)."""
    O = oracle_mod
    rng = np.random.default_rng(0)
    lvs = O.ompl_longest_segment(LO, HI, 0.002)
    path = np.loadtxt(os.path.join(REF, "alpha-1.2.org.path"))
    # motions that start on the solution path and end nearby: many are free
    # and many meet the environment somewhere inside
    s1 = path[rng.integers(0, len(path), 60)]
    s2 = s1.copy()
    s2[:, :3] += rng.normal(0, 3.0, (60, 3))
    rot = rng.standard_normal((60, 4)) * 0.15
    s2[:, 3:] += rot
    s2[:, 3:] /= np.linalg.norm(s2[:, 3:], axis=1, keepdims=True)
    results = []
    for a, b in zip(s1, s2):
        r, nd = alpha.ompl_check_motion(a, b, lvs)
        assert nd == math.ceil(O.ompl_distance(a, b) / lvs)
        want = alpha.ompl_is_valid(b) and all(alpha.ompl_is_valid(O.ompl_interpolate(a, b, k / nd))
                                              for k in range(1, nd))
        assert r == int(want)
        results.append(r)
    assert 10 < sum(results) < 50


def test_start_state_is_not_checked(alpha, oracle_mod):
    rng = np.random.default_rng(1)
    qs = random_ompl_states(200, rng)
    v = alpha.ompl_validate_states(qs)
    bad, good = qs[v == 0][0], qs[v == 1][0]
    # one segment: only the end state is tested
    lvs = oracle_mod.ompl_longest_segment(LO - 1e4, HI + 1e4, 0.5)
    assert oracle_mod.ompl_distance(bad, good) < lvs
    assert alpha.ompl_check_motion(bad, good, lvs) == (1, 1)
    assert alpha.ompl_check_motion(good, bad, lvs)[0] == 0
    # identical states: nd = 0, the end state alone decides
    assert alpha.ompl_check_motion(good, good, lvs) == (1, 0)
    assert alpha.ompl_check_motion(bad, bad, lvs) == (0, 0)


def test_alpha_solution_path_passes_the_motion_validator(alpha, oracle_mod, path):
    """res/alpha/alpha-1.2.org.path is an OMPL solution: every state is free
    and every consecutive motion passes the validator, down to the
    pipeline's default resolution 1e-4 (src/GP/pipeline/solve2.py:341);
    motions that skip ahead along the path do not."""
    assert alpha.ompl_validate_states(path).all()
    for res in (0.01, 1e-4):
        lvs = oracle_mod.ompl_longest_segment(LO, HI, res)
        out, sam = alpha.ompl_validate_motions(path[:-1], path[1:], lvs)
        assert out.all()
        assert sam.min() >= 1
    lvs = oracle_mod.ompl_longest_segment(LO, HI, 1e-3)
    out, _ = alpha.ompl_validate_motions(path[:-40], path[40:], lvs)
    assert out.mean() < 0.5


@pytest.mark.parametrize("center", [None, MOVED])
def test_oracle_translation_matches_library(oracle_mod, center, path):
    ow = oracle_mod.OracleWorld()
    ow.load_obj(0, ENV)
    ow.load_obj(1, ROB)
    uw = pyosr.UnitWorld()
    uw.loadModelFromFile(ENV)
    uw.loadRobotFromFile(ROB)
    if center is not None:
        ow.enforce_robot_center(center)
        uw.enforceRobotCenter(center)
    ow.scale_to_unit()
    ow.angle_model()
    uw.scaleToUnit()
    uw.angleModel(0.0, 0.0)
    assert np.array_equal(ow.ompl_center(), uw.get_ompl_center(1).astype(np.float32))
    qs = np.vstack([path, random_ompl_states(300, np.random.default_rng(2))])
    want = np.array([ow.translate_ompl_to_unit(q) for q in qs])
    assert np.array_equal(uw.translate_ompl_to_unit(qs), want)


def test_ompl_validity_does_not_depend_on_the_enforced_center(alpha, oracle_mod):
    """An OMPL state places the robot by its OMPL centre, whatever centre the
    unit world uses."""
    ow = oracle_mod.OracleWorld()
    ow.load_obj(0, ENV)
    ow.load_obj(1, ROB)
    ow.enforce_robot_center(MOVED)
    ow.scale_to_unit()
    ow.angle_model()
    qs = random_ompl_states(2000, np.random.default_rng(3))
    a = alpha.ompl_validate_states(qs)
    assert 0.05 < a.mean() < 0.95
    assert np.array_equal(a, ow.ompl_validate_states(qs))


def test_driver_host_rules(tmp_path):
    d = pyse3ompl.OmplDriver()
    with pytest.raises(RuntimeError):
        d.set_model_file(2, ENV)
    with pytest.raises(RuntimeError):
        d.set_state(2, [0, 0, 0], [1, 0, 0], 0.0)
    with pytest.raises(RuntimeError):
        d.validate_states(np.zeros((1, 7)))  # no geometry
    d.set_model_file(pyse3ompl.MODEL_PART_ENV, ENV)
    d.set_model_file(pyse3ompl.MODEL_PART_ROB, ROB)
    with pytest.raises(RuntimeError):
        d.validate_motion_pairs(np.zeros((1, 7)), np.zeros((1, 7)))  # no bounds
    d.set_bb(LO, HI)
    d.set_cdres(1e-4)
    assert d.longest_valid_segment == pytest.approx((np.linalg.norm(HI - LO) + math.pi / 2) * 1e-4, rel=1e-15)
    with pytest.raises(NotImplementedError):
        d.solve(0.01)
    with pytest.raises(ValueError):
        d.validate_states(np.zeros((2, 6)))
    assert pyse3ompl.PLANNER_RDT == pyse3ompl.PLANNER_ReRRT == 15


def write_alpha_cfg(tmp_path, with_res=True):
    import shutil
    for fn in (ENV, ROB):
        shutil.copy(fn, tmp_path)
    lines = ["[problem]", "name=Alpha-1.2", f"world={os.path.basename(ENV)}", f"robot={os.path.basename(ROB)}",
             "start.x=17.97", "start.y=7.23", "start.z=10.2", "start.axis.x=0", "start.axis.y=0", "start.axis.z=1",
             "start.theta=0", "goal.x=57.3", "goal.y=-2.5", "goal.z=18.4", "goal.axis.x=0.2",
             "goal.axis.y=0.4", "goal.axis.z=0.8", "goal.theta=1.5"]
    for k, lo, hi in zip("xyz", LO, HI):
        lines += [f"volume.min.{k}={float(lo)!r}", f"volume.max.{k}={float(hi)!r}"]
    if with_res:
        lines.append("collision_resolution=0.0001")
    fn = tmp_path / "alpha.cfg"
    fn.write_text("\n".join(lines) + "\n")
    return str(fn)


def test_puzzle_file(tmp_path):
    fn = write_alpha_cfg(tmp_path)
    cfg = pyse3ompl.read_puzzle(fn)
    assert cfg["env_fn"] == os.path.join(str(tmp_path), os.path.basename(ENV))
    assert np.array_equal(cfg["volume_min"], LO) and np.array_equal(cfg["volume_max"], HI)
    assert cfg["collision_resolution"] == 1e-4
    tr, axis, theta = cfg["goal"]
    assert np.array_equal(tr, [57.3, -2.5, 18.4]) and np.array_equal(axis, [0.2, 0.4, 0.8]) and theta == 1.5
    d = pyse3ompl.create_driver(fn)
    assert d.longest_valid_segment == pyse3ompl.longest_valid_segment(LO, HI, 1e-4)
    d = pyse3ompl.create_driver(fn, cdres=0.01, bvresize=5.0)
    assert d.longest_valid_segment == pyse3ompl.longest_valid_segment(LO - 5, HI + 5, 0.01)
    with pytest.raises(RuntimeError):
        pyse3ompl.create_driver(write_alpha_cfg(tmp_path, with_res=False))
    with pytest.raises(RuntimeError):
        pyse3ompl.read_puzzle(str(tmp_path / "missing.cfg"))
