"""ctypes binding of the CPU oracle (oracle/oracle.c).

Test infrastructure only: tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` test the GPU path against it and time
it as the CPU baseline.  The shipped ``puzzletunneldiscovery_amd`` package never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

_c_double_p = ctypes.POINTER(ctypes.c_double)
_c_float_p = ctypes.POINTER(ctypes.c_float)
_c_u32_p = ctypes.POINTER(ctypes.c_uint32)
_c_u8_p = ctypes.POINTER(ctypes.c_uint8)
_c_i8_p = ctypes.POINTER(ctypes.c_int8)
_c_i64_p = ctypes.POINTER(ctypes.c_int64)
_c_int_p = ctypes.POINTER(ctypes.c_int)

_lib = None


def build():
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        sig = {
            "or_set_libm": (None, [ctypes.c_int]),
            "or_get_libm": (ctypes.c_int, []),
            "or_world_new": (vp, []),
            "or_world_free": (None, [vp]),
            "or_load_arrays": (ctypes.c_int, [vp, ctypes.c_int, _c_float_p, ctypes.c_int64, _c_u32_p, ctypes.c_int64]),
            "or_load_obj": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_char_p]),
            "or_mesh_nv": (ctypes.c_int64, [vp, ctypes.c_int]),
            "or_mesh_nf": (ctypes.c_int64, [vp, ctypes.c_int]),
            "or_mesh_copy": (None, [vp, ctypes.c_int, _c_float_p, _c_u32_p]),
            "or_mesh_center": (None, [vp, ctypes.c_int, _c_float_p]),
            "or_enforce_robot_center": (None, [vp, _c_double_p]),
            "or_scale_to_unit": (None, [vp]),
            "or_angle_model": (ctypes.c_int, [vp, ctypes.c_float, ctypes.c_float]),
            "or_scene_scale": (ctypes.c_float, [vp]),
            "or_set_perturbation": (None, [vp, _c_double_p]),
            "or_cd_vertices": (None, [vp, ctypes.c_int, _c_double_p]),
            "or_calibration": (None, [vp, _c_double_p]),
            "or_bvh_depth": (ctypes.c_int, [vp, ctypes.c_int]),
            "or_state_to_transform": (None, [_c_double_p, _c_double_p, _c_double_p]),
            "or_distance": (ctypes.c_double, [_c_double_p, _c_double_p]),
            "or_interpolate": (None, [_c_double_p, _c_double_p, ctypes.c_double, _c_double_p]),
            "or_apply": (None, [_c_double_p, _c_double_p, _c_double_p, _c_double_p]),
            "or_translate_to_unit_state": (None, [vp, _c_double_p, _c_double_p]),
            "or_translate_ompl_to_unit": (None, [vp, _c_double_p, _c_double_p]),
            "or_ompl_center": (None, [vp, _c_float_p]),
            "or_ompl_distance": (ctypes.c_double, [_c_double_p, _c_double_p]),
            "or_ompl_longest_segment": (ctypes.c_double, [_c_double_p, _c_double_p, ctypes.c_double]),
            "or_ompl_interpolate": (None, [_c_double_p, _c_double_p, ctypes.c_double, _c_double_p]),
            "or_ompl_is_valid": (ctypes.c_int, [vp, _c_double_p]),
            "or_ompl_check_motion": (ctypes.c_int, [vp, _c_double_p, _c_double_p, ctypes.c_double, _c_i64_p]),
            "or_ompl_validate_states": (None, [vp, _c_double_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_int32),
                                               ctypes.c_int]),
            "or_ompl_validate_motions": (None, [vp, _c_double_p, _c_double_p, ctypes.c_int64, ctypes.c_double,
                                                ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                                ctypes.c_int]),
            "or_tri_intersect": (ctypes.c_int, [_c_double_p, _c_double_p]),
            "or_is_valid": (ctypes.c_int, [vp, _c_double_p]),
            "or_is_valid_counted": (ctypes.c_int, [vp, _c_double_p, _c_i64_p, _c_i64_p]),
            "or_is_valid_bruteforce": (ctypes.c_int, [vp, _c_double_p]),
            "or_valid_states": (None, [vp, _c_double_p, ctypes.c_int64, _c_u8_p, _c_i64_p, _c_i64_p, ctypes.c_int]),
            "or_transit": (None, [vp, _c_double_p, _c_double_p, ctypes.c_double, _c_double_p, _c_double_p,
                                  _c_int_p, _c_double_p, _c_double_p, _c_i64_p]),
            "or_is_valid_transition": (ctypes.c_int, [vp, _c_double_p, _c_double_p, ctypes.c_double, _c_i64_p]),
            "or_bisect": (None, [vp, _c_double_p, _c_double_p, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                 _c_double_p, _c_double_p]),
            "or_touch_batch": (None, [vp, _c_double_p, _c_double_p, ctypes.c_int64, ctypes.c_double, ctypes.c_int,
                                      _c_u8_p, _c_double_p, _c_double_p, _c_double_p, _c_double_p, _c_i64_p,
                                      ctypes.c_int]),
            "or_visibility_matrix": (None, [vp, _c_double_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double,
                                            _c_i8_p, ctypes.c_int]),
            "or_visibility_matrix2": (None, [vp, _c_double_p, ctypes.c_int64, ctypes.c_int, _c_double_p,
                                             ctypes.c_int64, ctypes.c_int, ctypes.c_double, _c_i8_p, ctypes.c_int]),
            "or_visibility_pair": (None, [vp, _c_double_p, _c_double_p, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_double, _c_i8_p, ctypes.c_int]),
            "or_contacts": (ctypes.c_int, [vp, _c_double_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int32),
                                           _c_double_p, _c_i8_p]),
            "or_tri_isect_segment": (ctypes.c_int, [_c_double_p, _c_double_p, _c_int_p, _c_double_p, _c_double_p]),
            "or_is_disentangled": (ctypes.c_int, [vp, _c_double_p]),
            "or_disentangled_states": (None, [vp, _c_double_p, ctypes.c_int64, _c_i8_p, ctypes.c_int]),
            "or_boxes_disjoint": (ctypes.c_int, [_c_double_p, _c_double_p, _c_double_p, _c_double_p, _c_double_p,
                                                 _c_double_p]),
            "or_model_box": (None, [vp, ctypes.c_int, _c_double_p, _c_double_p]),
            "or_se3_sin": (ctypes.c_double, [ctypes.c_double]),
            "or_se3_sin_legacy": (ctypes.c_double, [ctypes.c_double]),
            "or_cr_sin_slow": (ctypes.c_double, [ctypes.c_double]),
            "or_cr_cos_slow": (ctypes.c_double, [ctypes.c_double]),
            "or_cr_acos_slow": (ctypes.c_double, [ctypes.c_double]),
            "or_cr_atan2_slow": (ctypes.c_double, [ctypes.c_double, ctypes.c_double]),
            "or_se3_cos": (ctypes.c_double, [ctypes.c_double]),
            "or_se3_acos": (ctypes.c_double, [ctypes.c_double]),
            "or_se3_atan2": (ctypes.c_double, [ctypes.c_double, ctypes.c_double]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_c_double_p)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


def set_libm(mode):
    """0 or False: correctly rounded functions (the reference platform's
    libm); 1 or True: this host's glibc; 2: as 0, with the GPU kernels'
    faster sine for the states a walk only tests."""
    lib().or_set_libm(int(mode))


class OracleWorld:
    """UnitWorld restated on the CPU (binary64 FCL-style BVH)."""

    def __init__(self):
        self._w = lib().or_world_new()

    def __del__(self):
        w = getattr(self, "_w", None)
        if w and _lib is not None:
            _lib.or_world_free(w)
            self._w = None

    # ---- ingest -------------------------------------------------------
    def load_arrays(self, geo, V, F):
        V = np.ascontiguousarray(V, dtype=np.float32)
        F = np.ascontiguousarray(F, dtype=np.uint32)
        rc = lib().or_load_arrays(self._w, geo, V.ctypes.data_as(_c_float_p), V.shape[0],
                                  F.ctypes.data_as(_c_u32_p), F.shape[0])
        if rc != 0:
            raise RuntimeError(f"or_load_arrays failed ({rc})")

    def load_obj(self, geo, path):
        rc = lib().or_load_obj(self._w, geo, path.encode())
        if rc != 0:
            raise RuntimeError(f"or_load_obj failed ({rc})")

    def mesh(self, geo):
        nv = lib().or_mesh_nv(self._w, geo)
        nf = lib().or_mesh_nf(self._w, geo)
        V = np.zeros((nv, 3), np.float32)
        F = np.zeros((nf, 3), np.uint32)
        lib().or_mesh_copy(self._w, geo, V.ctypes.data_as(_c_float_p), F.ctypes.data_as(_c_u32_p))
        return V, F

    def center(self, geo):
        c = np.zeros(3, np.float32)
        lib().or_mesh_center(self._w, geo, c.ctypes.data_as(_c_float_p))
        return c

    def enforce_robot_center(self, c):
        c = _f64(c, (3,))
        lib().or_enforce_robot_center(self._w, _d(c))

    def scale_to_unit(self):
        lib().or_scale_to_unit(self._w)

    def angle_model(self, lat=0.0, lon=0.0):
        rc = lib().or_angle_model(self._w, lat, lon)
        if rc != 0:
            raise RuntimeError(f"or_angle_model failed ({rc})")

    def setup(self, env_V, env_F, rob_V, rob_F):
        self.load_arrays(0, env_V, env_F)
        self.load_arrays(1, rob_V, rob_F)
        self.scale_to_unit()
        self.angle_model()
        return self

    @property
    def scene_scale(self):
        return lib().or_scene_scale(self._w)

    def set_perturbation(self, q):
        q = _f64(q, (7,))
        lib().or_set_perturbation(self._w, _d(q))

    def cd_vertices(self, geo):
        nv = lib().or_mesh_nv(self._w, geo)
        out = np.zeros((nv, 3), np.float64)
        lib().or_cd_vertices(self._w, geo, _d(out))
        return out

    def calibration(self):
        m = np.zeros((4, 4), np.float64)
        lib().or_calibration(self._w, _d(m))
        return m

    def bvh_depth(self, geo):
        return lib().or_bvh_depth(self._w, geo)

    # ---- queries ------------------------------------------------------
    def is_valid(self, q):
        q = _f64(q, (7,))
        return bool(lib().or_is_valid(self._w, _d(q)))

    def is_valid_counted(self, q):
        q = _f64(q, (7,))
        a = ctypes.c_int64(0)
        b = ctypes.c_int64(0)
        v = lib().or_is_valid_counted(self._w, _d(q), ctypes.byref(a), ctypes.byref(b))
        return bool(v), a.value, b.value

    def is_valid_bruteforce(self, q):
        q = _f64(q, (7,))
        return bool(lib().or_is_valid_bruteforce(self._w, _d(q)))

    def valid_states(self, qs, nthreads=0, counters=False):
        qs = _f64(qs).reshape(-1, 7)
        n = qs.shape[0]
        out = np.zeros(n, np.uint8)
        nbv = np.zeros(n, np.int64)
        ntri = np.zeros(n, np.int64)
        lib().or_valid_states(self._w, _d(qs), n, out.ctypes.data_as(_c_u8_p),
                              nbv.ctypes.data_as(_c_i64_p), ntri.ctypes.data_as(_c_i64_p), nthreads)
        if counters:
            return out, nbv, ntri
        return out

    def is_disentangled(self, q):
        q = _f64(q, (7,))
        r = lib().or_is_disentangled(self._w, _d(q))
        if r < 0:
            raise RuntimeError("models not loaded")
        return bool(r)

    def disentangled_states(self, qs, nthreads=0):
        qs = _f64(qs).reshape(-1, 7)
        out = np.zeros(qs.shape[0], np.int8)
        lib().or_disentangled_states(self._w, _d(qs), qs.shape[0], out.ctypes.data_as(_c_i8_p), nthreads)
        if (out < 0).any():
            raise RuntimeError("models not loaded")
        return out.astype(bool)

    def contacts(self, q, cap=4096):
        """(pairs [m, 2], segs [m, 6], flags [m]) of all intersecting
        triangle pairs at q, (env, robot) order."""
        q = _f64(q, (7,))
        pairs = np.zeros((cap, 2), np.int32)
        segs = np.zeros((cap, 6))
        flags = np.zeros(cap, np.int8)
        m = lib().or_contacts(self._w, _d(q), cap, pairs.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                              _d(segs), flags.ctypes.data_as(_c_i8_p))
        if m < 0:
            raise RuntimeError("models not loaded")
        k = min(m, cap)
        return pairs[:k], segs[:k], flags[:k], m

    def model_box(self, geo):
        c = np.zeros(3)
        h = np.zeros(3)
        lib().or_model_box(self._w, geo, _d(c), _d(h))
        return c, h

    def transit(self, frm, to, vd):
        frm = _f64(frm, (7,))
        to = _f64(to, (7,))
        lf = np.zeros(7)
        st = np.zeros(7)
        done = ctypes.c_int(0)
        lt = ctypes.c_double(0)
        t = ctypes.c_double(0)
        ns = ctypes.c_int64(0)
        lib().or_transit(self._w, _d(frm), _d(to), vd, _d(lf), _d(st), ctypes.byref(done),
                         ctypes.byref(lt), ctypes.byref(t), ctypes.byref(ns))
        return lf, st, bool(done.value), lt.value, t.value, ns.value

    def is_valid_transition(self, frm, to, ivd, with_steps=False):
        frm = _f64(frm, (7,))
        to = _f64(to, (7,))
        ns = ctypes.c_int64(0)
        ok = bool(lib().or_is_valid_transition(self._w, _d(frm), _d(to), ivd, ctypes.byref(ns)))
        return (ok, ns.value) if with_steps else ok

    def bisect(self, frm, to, lo, hi, levels):
        frm = _f64(frm, (7,))
        to = _f64(to, (7,))
        a = ctypes.c_double(0)
        b = ctypes.c_double(0)
        lib().or_bisect(self._w, _d(frm), _d(to), lo, hi, levels, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def touch_batch(self, frm, to, vd, levels, nthreads=0):
        frm = _f64(frm).reshape(-1, 7)
        to = _f64(to).reshape(-1, 7)
        n = frm.shape[0]
        done = np.zeros(n, np.uint8)
        lt = np.zeros(n)
        t = np.zeros(n)
        lo = np.zeros(n)
        hi = np.zeros(n)
        q = np.zeros(n, np.int64)
        lib().or_touch_batch(self._w, _d(frm), _d(to), n, vd, levels, done.ctypes.data_as(_c_u8_p),
                             _d(lt), _d(t), _d(lo), _d(hi), q.ctypes.data_as(_c_i64_p), nthreads)
        return dict(done=done, last_tau=lt, tau=t, lo=lo, hi=hi, queries=q)

    def translate_to_unit_state(self, q):
        q = _f64(q, (7,))
        out = np.zeros(7)
        lib().or_translate_to_unit_state(self._w, _d(q), _d(out))
        return out

    def translate_ompl_to_unit(self, q):
        q = _f64(q, (7,))
        out = np.zeros(7)
        lib().or_translate_ompl_to_unit(self._w, _d(q), _d(out))
        return out

    def ompl_center(self):
        c = np.zeros(3, np.float32)
        lib().or_ompl_center(self._w, c.ctypes.data_as(_c_float_p))
        return c

    def ompl_is_valid(self, q):
        return bool(lib().or_ompl_is_valid(self._w, _d(_f64(q, (7,)))))

    def ompl_check_motion(self, s1, s2, lvs):
        """(result, nd): DiscreteMotionValidator::checkMotion as OMPL runs it."""
        nd = ctypes.c_int64(0)
        r = lib().or_ompl_check_motion(self._w, _d(_f64(s1, (7,))), _d(_f64(s2, (7,))), float(lvs),
                                       ctypes.byref(nd))
        return r, nd.value

    def ompl_validate_states(self, qs, nthreads=0):
        qs = _f64(qs).reshape(-1, 7)
        out = np.zeros(len(qs), np.int32)
        lib().or_ompl_validate_states(self._w, _d(qs), len(qs),
                                      out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), nthreads)
        return out

    def ompl_validate_motions(self, q0, q1, lvs, nthreads=0):
        """(out, samples) for the motions q0[i] -> q1[i]."""
        q0 = _f64(q0).reshape(-1, 7)
        q1 = _f64(q1).reshape(-1, 7)
        n = min(len(q0), len(q1))
        out = np.zeros(n, np.int32)
        sam = np.zeros(n, np.int32)
        lib().or_ompl_validate_motions(self._w, _d(q0), _d(q1), n, float(lvs),
                                       out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                       sam.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), nthreads)
        return out, sam

    def visibility_matrix(self, qs, is_unit, vd, nthreads=0):
        qs = _f64(qs).reshape(-1, 7)
        n = qs.shape[0]
        out = np.zeros((n, n), np.int8)
        lib().or_visibility_matrix(self._w, _d(qs), n, int(is_unit), vd, out.ctypes.data_as(_c_i8_p), nthreads)
        return out

    def visibility_matrix2(self, qs0, unit0, qs1, unit1, vd, nthreads=0):
        qs0 = _f64(qs0).reshape(-1, 7)
        qs1 = _f64(qs1).reshape(-1, 7)
        out = np.zeros((qs0.shape[0], qs1.shape[0]), np.int8)
        lib().or_visibility_matrix2(self._w, _d(qs0), qs0.shape[0], int(unit0), _d(qs1), qs1.shape[0],
                                    int(unit1), vd, out.ctypes.data_as(_c_i8_p), nthreads)
        return out

    def visibility_pair(self, qs0, unit0, qs1, unit1, vd, nthreads=0):
        qs0 = _f64(qs0).reshape(-1, 7)
        qs1 = _f64(qs1).reshape(-1, 7)
        if qs0.shape[0] != qs1.shape[0]:
            raise ValueError("visibility_pair needs equal row counts")
        out = np.zeros(qs0.shape[0], np.int8)
        lib().or_visibility_pair(self._w, _d(qs0), _d(qs1), qs0.shape[0], int(unit0), int(unit1), vd,
                                 out.ctypes.data_as(_c_i8_p), nthreads)
        return out


def distance(a, b):
    a = _f64(a, (7,))
    b = _f64(b, (7,))
    return lib().or_distance(_d(a), _d(b))


def interpolate(a, b, tau):
    a = _f64(a, (7,))
    b = _f64(b, (7,))
    out = np.zeros(7)
    lib().or_interpolate(_d(a), _d(b), tau, _d(out))
    return out


def apply(frm, tr, aa):
    frm = _f64(frm, (7,))
    tr = _f64(tr, (3,))
    aa = _f64(aa, (3,))
    out = np.zeros(7)
    lib().or_apply(_d(frm), _d(tr), _d(aa), _d(out))
    return out


def state_to_transform(q):
    q = _f64(q, (7,))
    R = np.zeros((3, 3))
    t = np.zeros(3)
    lib().or_state_to_transform(_d(q), _d(R), _d(t))
    return R, t


def ompl_distance(a, b):
    return lib().or_ompl_distance(_d(_f64(a, (7,))), _d(_f64(b, (7,))))


def ompl_longest_segment(lo, hi, resolution):
    return lib().or_ompl_longest_segment(_d(_f64(lo, (3,))), _d(_f64(hi, (3,))), float(resolution))


def ompl_interpolate(a, b, t):
    out = np.zeros(7)
    lib().or_ompl_interpolate(_d(_f64(a, (7,))), _d(_f64(b, (7,))), float(t), _d(out))
    return out


def tri_isect_segment(V, U):
    """Moller's triangle-triangle test with the intersection segment:
    (meets, coplanar, p1, p2)."""
    V = _f64(V, (9,))
    U = _f64(U, (9,))
    cop = ctypes.c_int(0)
    p1 = np.zeros(3)
    p2 = np.zeros(3)
    r = lib().or_tri_isect_segment(_d(V), _d(U), ctypes.byref(cop), _d(p1), _d(p2))
    return bool(r), bool(cop.value), p1, p2


def boxes_disjoint(R1, p1, A, R2, p2, B):
    """FCL's box-box test (boolean): 1 if the boxes are disjoint."""
    a = [_f64(x).reshape(-1) for x in (R1, p1, A, R2, p2, B)]
    return bool(lib().or_boxes_disjoint(*[_d(x) for x in a]))


def tri_intersect(P, Q):
    P = _f64(P, (9,))
    Q = _f64(Q, (9,))
    return bool(lib().or_tri_intersect(_d(P), _d(Q)))
