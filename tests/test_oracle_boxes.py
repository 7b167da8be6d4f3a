"""Pins the oracle's isDisentangled (unit_world.cc:249-264 -> cdmodel.cc:180-198:
FCL box shapes of both models' AABBs and FCL's box-box test) against closed
forms and against an independent exact decision: two boxes intersect iff the
linear program "find a point inside both" is feasible (scipy's HiGHS)."""
import numpy as np
import pytest
from scipy.optimize import linprog
from scipy.spatial.transform import Rotation

from puzzletunneldiscovery_amd import workloads as W


def lp_margin(R1, p1, A, R2, p2, B):
    """max t such that some x has |Ri^T (x - pi)|_k <= hi_k - t for both
    boxes: > 0 means the interiors meet, < 0 means the boxes are disjoint."""
    rows, rhs = [], []
    for R, p, h in ((R1, p1, A), (R2, p2, B)):
        for k in range(3):
            ax = R[:, k]
            for sgn in (1.0, -1.0):
                rows.append(np.concatenate([sgn * ax, [1.0]]))
                rhs.append(h[k] + sgn * ax @ p)
    res = linprog(c=[0, 0, 0, -1.0], A_ub=np.array(rows), b_ub=np.array(rhs),
                  bounds=[(None, None)] * 3 + [(None, 10.0)], method="highs")
    assert res.status == 0
    return -res.fun


def test_axis_aligned_closed_form(oracle_mod):
    I = np.eye(3)
    h = np.ones(3)
    for gap, disjoint in ((1e-9, True), (0.0, False), (-0.5, False), (0.3, True)):
        assert oracle_mod.boxes_disjoint(I, [0, 0, 0], h, I, [2 + gap, 0, 0], h) == disjoint
        assert oracle_mod.boxes_disjoint(I, [0, 0, 0], h, I, [0, -(2 + gap), 0], h) == disjoint


def test_rotated_closed_form(oracle_mod):
    # box 2 turned 45 degrees about z reaches sqrt(2) along x
    R2 = Rotation.from_euler("z", 45, degrees=True).as_matrix()
    h = np.ones(3)
    edge = 1 + np.sqrt(2)
    assert oracle_mod.boxes_disjoint(np.eye(3), [0, 0, 0], h, R2, [edge + 1e-6, 0, 0], h)
    assert not oracle_mod.boxes_disjoint(np.eye(3), [0, 0, 0], h, R2, [edge - 1e-6, 0, 0], h)


def test_random_boxes_match_lp(oracle_mod):
    rng = np.random.default_rng(3)
    decided = both = 0
    for _ in range(1500):
        R1 = Rotation.random(random_state=rng).as_matrix()
        R2 = Rotation.random(random_state=rng).as_matrix()
        A = rng.uniform(0.05, 1.0, 3)
        B = rng.uniform(0.05, 1.0, 3)
        p1 = rng.uniform(-1, 1, 3)
        p2 = rng.uniform(-1, 1, 3) * 1.5
        t = lp_margin(R1, p1, A, R2, p2, B)
        if abs(t) < 1e-5:
            continue  # the 1e-6 edge guard may decide either way this close
        got = oracle_mod.boxes_disjoint(R1, p1, A, R2, p2, B)
        assert got == (t < 0), (t, got)
        decided += 1
        both += got
    assert decided > 1400 and 200 < both < decided - 200


def test_world_disentangled(oracle_mod, torus_oracle):
    far = np.array([5.0, 5.0, 5.0, 1, 0, 0, 0])
    assert torus_oracle.is_disentangled(far)
    centre, _ = torus_oracle.model_box(0)
    at = np.concatenate([centre, [1, 0, 0, 0]])
    assert not torus_oracle.is_disentangled(at)
    qs = W.random_states(2000, seed=4, lo=-0.2, hi=1.2)
    d = torus_oracle.disentangled_states(qs)
    assert 0.1 < d.mean() < 0.9
    # a disentangled state has no colliding triangle pair
    v = torus_oracle.valid_states(qs).astype(bool)
    assert v[d].all()
    with pytest.raises(RuntimeError):
        oracle_mod.OracleWorld().is_disentangled(far)
