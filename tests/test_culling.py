"""The binary32 culling test of the traversals (csrc/cull.h: the culling
transform make_qcull, the robot box placed by load_rob_box, the 15-axis
obb_disjoint) against exact separation.  The exact leaf test maps a robot
point v to Re^T (R(q) v + t - te), R(q) the matrix of the raw (possibly
non-unit) state quaternion and Re that of the raw perturbation quaternion
(kernels.hip make_exact).  Every box pair the binary32 test culls must be
disjoint under that map in exact rational arithmetic, decided by the
separating-axis theorem over the two parallelepipeds' 3 + 3 face normals and
9 edge-pair normals.  Pairs are built near touching (gaps from 1e-9 to 0.1
of the boxes' summed extents, separated and overlapping, so that culls
happen on both sides of the margins), with non-unit quaternions, and with
perturbations that move the environment up to 1,000 units away.

The culling code is built for the host here (g++, contraction off); the
device build differs only in rsqrtf, whose error the quaternion-norm term
of the inflation covers.  With the margins removed, the same pairs must
expose culls of overlapping boxes, so the check is known to be able to
fail."""
import ctypes
import math
import os
import subprocess
from fractions import Fraction as Fr

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cullcheck", "cullcheck.cpp")


@pytest.fixture(scope="module")
def cc(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("cullcheck") / "libcullcheck.so")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", SRC, "-o", out])
    lib = ctypes.CDLL(out)
    fp = ctypes.POINTER(ctypes.c_float)
    dp = ctypes.POINTER(ctypes.c_double)
    lib.cc_culled.argtypes = [fp, fp, dp, dp, ctypes.c_float, ctypes.c_float, ctypes.c_int]
    lib.cc_culled.restype = ctypes.c_int
    return lib


def _rot(w, x, y, z):
    """se3_quat_to_rotmat's polynomial (exact for Fractions, else float)."""
    tx, ty, tz = 2 * x, 2 * y, 2 * z
    twx, twy, twz = tx * w, ty * w, tz * w
    txx, txy, txz = tx * x, ty * x, tz * x
    tyy, tyz, tzz = ty * y, tz * y, tz * z
    return [[1 - (tyy + tzz), txy - twz, txz + twy],
            [txy + twz, 1 - (txx + tzz), tyz - twx],
            [txz - twy, tyz + twx, 1 - (txx + tyy)]]


def _mv(M, v):
    return [M[i][0] * v[0] + M[i][1] * v[1] + M[i][2] * v[2] for i in range(3)]


def _mtv(M, v):
    return [M[0][i] * v[0] + M[1][i] * v[1] + M[2][i] * v[2] for i in range(3)]


def _cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def _dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def _placed(env, rob, q, pert, num):
    """Centres and half-edge vectors of the environment box and of the robot
    box under the exact map, in number type num (Fraction or float)."""
    cA = [num(float(v)) for v in env[0:3]]
    EA = [[num(float(env[4 * (i + 1) + k])) * num(float(env[4 * i + 3])) for k in range(3)] for i in range(3)]
    cB = [num(float(v)) for v in rob[0:3]]
    EB = [[num(float(rob[4 * (i + 1) + k])) * num(float(rob[4 * i + 3])) for k in range(3)] for i in range(3)]
    R = _rot(*[num(float(v)) for v in q[3:7]])
    t = [num(float(v)) for v in q[0:3]]
    if pert is None:
        Re, te = [[num(1), num(0), num(0)], [num(0), num(1), num(0)], [num(0), num(0), num(1)]], [num(0)] * 3
    else:
        Re, te = _rot(*[num(float(v)) for v in pert[3:7]]), [num(float(v)) for v in pert[0:3]]
    Rc = _mv(R, cB)
    cBp = _mtv(Re, [Rc[k] + t[k] - te[k] for k in range(3)])
    EBp = [_mtv(Re, _mv(R, e)) for e in EB]
    return cA, EA, cBp, EBp


def _gap(cA, EA, cB, EB):
    """max over the 15 axes of (|d.n| - rA - rB) / |n| (> 0: separated)."""
    d = [cB[k] - cA[k] for k in range(3)]
    axes = [_cross(EA[1], EA[2]), _cross(EA[2], EA[0]), _cross(EA[0], EA[1]),
            _cross(EB[1], EB[2]), _cross(EB[2], EB[0]), _cross(EB[0], EB[1])]
    axes += [_cross(a, b) for a in EA for b in EB]
    best = -math.inf
    for n in axes:
        nn = _dot(n, n)
        if nn == 0:
            continue
        r = sum(abs(_dot(e, n)) for e in EA) + sum(abs(_dot(e, n)) for e in EB)
        g = abs(_dot(d, n)) - r
        if isinstance(g, Fr):
            if g > 0:
                return 1.0
            continue
        best = max(best, g / math.sqrt(nn))
    return best if not isinstance(cA[0], Fr) else -1.0


def _node(rng, centre):
    """A GNode's 16 floats: centre, extents, orthonormal-ish binary32 axes."""
    qq = rng.normal(size=4)
    qq /= np.linalg.norm(qq)
    R = np.array(_rot(*qq))
    e = 10.0 ** rng.uniform(-3.0, -0.5, 3)
    v = np.zeros(16, np.float32)
    v[0:3] = centre
    v[3], v[7], v[11] = e
    v[4:7], v[8:11], v[12:15] = R[0], R[1], R[2]
    v[15] = np.array([-1], np.int32).view(np.float32)[0]
    return v


def _case(rng, kind):
    env = _node(rng, rng.uniform(-0.5, 0.5, 3))
    rob = _node(rng, rng.uniform(-0.1, 0.1, 3))
    qq = rng.normal(size=4)
    qq /= np.linalg.norm(qq)
    if kind == "unit":
        qn = qq
    elif kind == "nonunit":
        qn = qq * rng.uniform(0.7, 1.3)
    else:
        qn = qq * (1.0 + rng.uniform(-3e-7, 3e-7))
    pert = None
    if rng.random() < 0.5:
        pq = rng.normal(size=4)
        pq /= np.linalg.norm(pq)
        if rng.random() < 0.5:
            pq *= rng.uniform(0.9, 1.1)
        te = rng.uniform(-1.0, 1.0, 3) * 10.0 ** rng.uniform(0, 3)
        pert = np.concatenate([te, pq])
    # place the robot box's centre at a point of a ray from the env box's
    # centre, moved along it until the exact-ish gap is the wanted one
    Rm = np.array(_rot(*qn))
    Re = np.eye(3) if pert is None else np.array(_rot(*pert[3:]))
    te = np.zeros(3) if pert is None else pert[:3]
    u = rng.normal(size=3)
    u /= np.linalg.norm(u)
    scale = float(np.sum(env[[3, 7, 11]]) + np.sum(rob[[3, 7, 11]]))
    want = (1.0 if rng.random() < 0.7 else -1.0) * scale * 10.0 ** rng.uniform(-9, -1)

    def state(s):
        target = env[0:3].astype(np.float64) + s * u
        # Re^T (R c + t - te) = target  ->  t = Re^-T target - R c + te
        t = np.linalg.solve(Re.T, target) - Rm @ rob[0:3].astype(np.float64) + te
        return np.concatenate([t, qn])

    def gap(s):
        return _gap(*_placed(env, rob, state(s), pert, float))

    lo, hi = 0.0, 4.0 * scale + 1.0
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        if gap(mid) < want:
            lo = mid
        else:
            hi = mid
    return env, rob, state(hi), pert


def _run(cc, margins, n, seed):
    rng = np.random.default_rng(seed)
    fp = ctypes.POINTER(ctypes.c_float)
    dp = ctypes.POINTER(ctypes.c_double)
    culled = wrong = near_culled = 0
    for i in range(n):
        env, rob, q, pert = _case(rng, ("unit", "nonunit", "nearly")[i % 3])
        reach = np.float32(np.linalg.norm(rob[0:3].astype(np.float64)) +
                           np.linalg.norm(rob[[3, 7, 11]].astype(np.float64)))
        reach = np.nextafter(reach, np.float32(np.inf))
        mag = np.nextafter(np.float32(np.abs(env[0:3]).max() + env[[3, 7, 11]].max() + np.abs(rob[0:3]).max() +
                                      rob[[3, 7, 11]].max()), np.float32(np.inf))
        q = np.ascontiguousarray(q, np.float64)
        p = None if pert is None else np.ascontiguousarray(pert, np.float64)
        c = cc.cc_culled(env.ctypes.data_as(fp), rob.ctypes.data_as(fp), q.ctypes.data_as(dp),
                         None if p is None else p.ctypes.data_as(dp), float(reach), float(mag), margins)
        if not c:
            continue
        culled += 1
        g = _gap(*_placed(env, rob, q, pert, float))
        scale = float(np.sum(env[[3, 7, 11]]) + np.sum(rob[[3, 7, 11]]))
        if g > 1e-7 * scale:
            continue  # separated by far more than binary64 rounding
        near_culled += 1
        if _gap(*_placed(env, rob, q, pert, Fr)) <= 0:
            wrong += 1
    return culled, near_culled, wrong


def test_culled_pairs_are_disjoint(cc):
    culled, near, wrong = _run(cc, 1, 1500, 11)
    assert culled > 150
    assert wrong == 0


def test_without_margins_overlaps_are_culled(cc):
    culled, near, wrong = _run(cc, 0, 1500, 11)
    assert wrong > 0
