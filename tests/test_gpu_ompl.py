"""OMPL-side validity on the GPU (pyse3ompl.OmplDriver.validate_states and
validate_motion_pairs, lib/pyse3ompl/ompldriver.cc:566-620) against the
oracle's restatement of OMPL's discrete motion validator: decisions and
sample counts equal, on the alpha puzzle (the reference's own meshes and
OMPL solution path) and on the torus scene."""
import math
import os

import numpy as np
import pytest

from puzzletunneldiscovery_amd import pyosr, pyse3ompl

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(HERE, "golden", "ref")
ENV = os.path.join(REF, "alpha_env-1.2.org.obj")
ROB = os.path.join(REF, "alpha-1.2.org.obj")
MOVED = np.array([3.5, -7.25, 12.0])
LO = np.array([-11.892796, -13.170522, -15.239033])
HI = np.array([68.73492, 75.37971186, 56.18796844])


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(scope="module")
def driver(torch_cuda):
    d = pyse3ompl.OmplDriver()
    d.set_model_file(pyse3ompl.MODEL_PART_ENV, ENV)
    d.set_model_file(pyse3ompl.MODEL_PART_ROB, ROB)
    d.set_bb(LO, HI)
    d.set_cdres(1e-4)
    return d


@pytest.fixture(scope="module")
def alpha(oracle_mod):
    ow = oracle_mod.OracleWorld()
    ow.load_obj(0, ENV)
    ow.load_obj(1, ROB)
    ow.scale_to_unit()
    ow.angle_model()
    return ow


def random_ompl_states(n, rng, lo=LO + 10, hi=HI - 10):
    q = np.zeros((n, 7))
    q[:, :3] = rng.uniform(lo, hi, (n, 3))
    r = rng.standard_normal((n, 4))
    q[:, 3:] = r / np.linalg.norm(r, axis=1, keepdims=True)
    return q


def test_validate_states_matches_oracle(driver, alpha):
    path = np.loadtxt(os.path.join(REF, "alpha-1.2.org.path"))
    assert driver.validate_states(path).all()
    qs = random_ompl_states(20000, np.random.default_rng(10))
    got = driver.validate_states(qs)
    assert got.dtype == np.int32 and got.shape == (20000,)
    want = alpha.ompl_validate_states(qs)
    assert 0.05 < want.mean() < 0.95
    assert np.array_equal(got, want)
    assert driver.validate_states(np.zeros((0, 7))).shape == (0,)


def test_validate_states_with_moved_center(torch_cuda, alpha):
    uw = pyosr.UnitWorld()
    uw.loadModelFromFile(ENV)
    uw.loadRobotFromFile(ROB)
    uw.enforceRobotCenter(MOVED)
    uw.scaleToUnit()
    uw.angleModel(0.0, 0.0)
    d = pyse3ompl.OmplDriver.from_unit_world(uw)
    qs = random_ompl_states(5000, np.random.default_rng(11))
    assert np.array_equal(d.validate_states(qs), alpha.ompl_validate_states(qs))


def edge_case_motions(path, rng):
    """Identical states (nd = 0), the same rotation written as -q (long-arc
    sign), rotations below OMPL's arc threshold, and short hops."""
    a = path[rng.integers(0, len(path), 40)]
    b = a.copy()
    b[:10, 3:] = -b[:10, 3:]
    b[10:20, :3] += 0.05
    b[20:30, 3:] += 1e-6 * rng.standard_normal((10, 4))
    b[20:30, 3:] /= np.linalg.norm(b[20:30, 3:], axis=1, keepdims=True)
    b[30:, :3] += rng.normal(0, 1.0, (10, 3))
    return a, b


def test_validate_motion_pairs_matches_oracle(driver, alpha):
    path = np.loadtxt(os.path.join(REF, "alpha-1.2.org.path"))
    lvs = driver.longest_valid_segment
    # the solution path: every consecutive motion is valid
    out, sam = driver.validate_motion_pairs(path[:-1], path[1:], return_samples=True)
    want, wsam = alpha.ompl_validate_motions(path[:-1], path[1:], lvs)
    assert out.all() and np.array_equal(out, want) and np.array_equal(sam, wsam)
    rng = np.random.default_rng(12)
    # key-configuration screening (src/GP/pipeline/solve2.py:219): all pairs
    # of free states, here the path states and free random states
    qs = random_ompl_states(3000, rng)
    keys = np.vstack([path[::4], qs[alpha.ompl_validate_states(qs) == 1][:40]])
    fi, ti = np.tril_indices(len(keys))
    a, b = edge_case_motions(path, rng)
    q0 = np.vstack([keys[fi], a])
    q1 = np.vstack([keys[ti], b])
    out, sam = driver.validate_motion_pairs(q0, q1, return_samples=True)
    want, wsam = alpha.ompl_validate_motions(q0, q1, lvs)
    assert np.array_equal(sam, wsam)
    assert np.array_equal(out, want)
    assert 0.02 < out.mean() < 0.9, out.mean()
    # unequal row counts: the shorter one decides
    assert driver.validate_motion_pairs(q0[:5], q1[:3]).shape == (3,)


def test_motion_device_form(driver, alpha, torch_cuda):
    torch = torch_cuda
    rng = np.random.default_rng(13)
    path = np.loadtxt(os.path.join(REF, "alpha-1.2.org.path"))
    q0 = np.vstack([path[:-1], random_ompl_states(200, rng)])
    q1 = np.vstack([path[1:], random_ompl_states(200, rng)])
    host, hsam = driver.validate_motion_pairs(q0, q1, return_samples=True)
    d0 = torch.from_numpy(q0).cuda()
    d1 = torch.from_numpy(q1).cuda()
    out = torch.full((len(q0),), 7, dtype=torch.int32, device="cuda")
    sam = torch.zeros(len(q0), dtype=torch.int32, device="cuda")
    s = torch.cuda.current_stream()
    driver.validate_motion_pairs_device(d0.data_ptr(), d1.data_ptr(), len(q0), out.data_ptr(), sam.data_ptr(),
                                        stream=s.cuda_stream)
    st = torch.zeros(len(q0), dtype=torch.int32, device="cuda")
    driver.validate_states_device(d1.data_ptr(), len(q1), st.data_ptr(), stream=s.cuda_stream)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), host)
    assert np.array_equal(sam.cpu().numpy(), hsam)
    assert np.array_equal(st.cpu().numpy(), driver.validate_states(q1))
    # a non-finite state: the device form flags it, the host form raises
    bad = q1.copy()
    bad[0, 0] = np.inf
    with pytest.raises(RuntimeError):
        driver.validate_motion_pairs(q0, bad)


def test_torus_motions(gpu_world, torus_oracle):
    d = pyse3ompl.OmplDriver.from_unit_world(gpu_world)
    d.set_bb([-1, -1, -1], [1, 1, 1])
    d.set_cdres(0.001)
    rng = np.random.default_rng(14)
    n = 2000
    q0 = np.zeros((n, 7))
    q0[:, :3] = rng.uniform(0.15, 0.85, (n, 3))
    r = rng.standard_normal((n, 4))
    q0[:, 3:] = r / np.linalg.norm(r, axis=1, keepdims=True)
    q1 = q0.copy()
    q1[:, :3] += rng.normal(0, 0.15, (n, 3))
    r = q1[:, 3:] + 0.3 * rng.standard_normal((n, 4))
    q1[:, 3:] = r / np.linalg.norm(r, axis=1, keepdims=True)
    out, sam = d.validate_motion_pairs(q0, q1, return_samples=True)
    want, wsam = torus_oracle.ompl_validate_motions(q0, q1, d.longest_valid_segment)
    assert np.array_equal(sam, wsam)
    assert np.array_equal(out, want)
    assert 0.05 < out.mean() < 0.95
    assert math.isclose(d.longest_valid_segment, (math.sqrt(12) + math.pi / 2) * 0.001, rel_tol=1e-15)
