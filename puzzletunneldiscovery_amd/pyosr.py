"""pyosr-compatible Python surface over the gfx950 collision-query engine.

Drop-in for the validity part of the reference's ``pyosr`` module
(lib/pyosr/module.cc:77-125,127-250,328-333): the same class and method
names, argument names, defaults and return shapes, so callers such as
src/GP/pipeline/uw_random.py, touchq_util.py and solve2.py can switch with
``from puzzletunneldiscovery_amd import pyosr``.

Beyond the reference's one-state-per-call methods, ``UnitWorld`` offers
batched forms (``is_valid_states``, ``transit_states_to_with_contact``,
``is_valid_transitions``, ``touch_search``) that hand a whole array to the GPU
in one call; the per-state methods are thin wrappers over them.

States are rows [tx, ty, tz, qw, qx, qy, qz]; binary32 inputs are widened to
binary64 as pybind11's Eigen casters do.  Taus in transit tuples are narrowed
to binary32 like the reference's ``std::tuple<..., float, float>``.
"""
import ctypes

import numpy as np

from . import _native as N

STATE_DIMENSION = 7


def _state(q):
    q = np.ascontiguousarray(np.asarray(q, dtype=np.float64).reshape(-1))
    if q.shape[0] != 7:
        raise ValueError("a state has 7 elements [tx, ty, tz, qw, qx, qy, qz]")
    return q


def _states(qs):
    qs = np.ascontiguousarray(np.asarray(qs, dtype=np.float64))
    if qs.ndim == 1:
        qs = qs.reshape(1, -1)
    if qs.shape[-1] != 7:
        raise ValueError("states must have 7 columns")
    return np.ascontiguousarray(qs.reshape(-1, 7))


def _vec3(v):
    v = np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(-1))
    if v.shape[0] != 3:
        raise ValueError("expected 3 elements")
    return v


def get_identity_state_vector():
    return np.array([0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0])


def distance(lhv, rhv):
    """distance (lib/osr/osr_state.cc:130-154)."""
    a, b = _state(lhv), _state(rhv)
    out = ctypes.c_double(0.0)
    N.check(N.lib().se3cq_distance(N.dptr(a), N.dptr(b), ctypes.byref(out)))
    return out.value


def multi_distance(origin, targets):
    """multi_distance (osr_state.cc:156-166)."""
    o = _state(origin)
    t = _states(targets)
    out = np.zeros(t.shape[0])
    N.check(N.lib().se3cq_multi_distance(N.dptr(o), N.dptr(t), t.shape[0], N.dptr(out)))
    return out


def interpolate(pkey, nkey, tau):
    """interpolate (osr_state.cc:27-44): lerp translation, slerp rotation."""
    a, b = _state(pkey), _state(nkey)
    out = np.zeros(7)
    N.check(N.lib().se3cq_interpolate(N.dptr(a), N.dptr(b), float(tau), N.dptr(out)))
    return out


def apply(frm, tr, aa):
    """apply (osr_state.cc:243-259)."""
    f, t, a = _state(frm), _vec3(tr), _vec3(aa)
    out = np.zeros(7)
    N.check(N.lib().se3cq_apply(N.dptr(f), N.dptr(t), N.dptr(a), N.dptr(out)))
    return out


def apply_batch(frms, trs, aas):
    """apply over rows: [n, 7] states moved by [n, 3] translations and
    [n, 3] axis-angle rotations (the pipeline's touch sampling,
    src/GP/pipeline/touchq_util.py:8-13, calls apply once per sample)."""
    f = _states(frms)
    t = np.ascontiguousarray(trs, dtype=np.float64).reshape(-1, 3)
    a = np.ascontiguousarray(aas, dtype=np.float64).reshape(-1, 3)
    if not (len(f) == len(t) == len(a)):
        raise ValueError("apply_batch: row counts differ")
    out = np.zeros((len(f), 7))
    N.check(N.lib().se3cq_apply_batch(N.dptr(f), N.dptr(t), N.dptr(a), len(f), N.dptr(out)))
    return out


def decompose_2(state):
    """decompose_2 (osr_state.cc:94-99): (translation, 3x3 rotation)."""
    q = _state(state)
    t = np.zeros(3)
    R = np.zeros((3, 3))
    N.check(N.lib().se3cq_decompose_2(N.dptr(q), N.dptr(t), N.dptr(R)))
    return t, R


def decompose_3(state):
    """decompose_3 (osr_state.cc:101-107): (translation, angle, axis) of the
    normalised state rotation."""
    q = _state(state)
    t = np.zeros(3)
    a = np.zeros(1)
    axis = np.zeros(3)
    N.check(N.lib().se3cq_decompose_3(N.dptr(q), N.dptr(t), N.dptr(a), N.dptr(axis)))
    return t, float(a[0]), axis


def compose_from_angleaxis(tr, angle, axis):
    """compose_from_angleaxis (osr_state.cc:121-128)."""
    t, a = _vec3(tr), _vec3(axis)
    out = np.zeros(7)
    N.check(N.lib().se3cq_compose_from_angleaxis(N.dptr(t), float(angle), N.dptr(a), N.dptr(out)))
    return out


class UnitWorld:
    """Unit-world collision queries on the GPU (osr::UnitWorld's validity API)."""

    GEO_ENV = 0
    GEO_ROB = 1

    def __init__(self):
        L = N.lib()
        h = ctypes.c_void_p()
        N.check(L.se3cq_world_create(ctypes.byref(h)))
        self._w = h
        self._state = get_identity_state_vector()

    def __del__(self):
        h = getattr(self, "_w", None)
        if h and N is not None and N._lib is not None:
            N._lib.se3cq_world_destroy(h)
            self._w = None

    # ---- setup (module.cc:129-137) ----------------------------------
    def setupFrom(self, other):
        N.check(N.lib().se3cq_world_copy_from(self._w, other._w))

    def loadModelFromFile(self, fn):
        N.check(N.lib().se3cq_load_file(self._w, 0, str(fn).encode()))

    def loadRobotFromFile(self, fn):
        N.check(N.lib().se3cq_load_file(self._w, 1, str(fn).encode()))
        self._state = get_identity_state_vector()

    def load_mesh(self, geo, V, F):
        """Load a mesh from binary32 vertices [nv,3] and triangles [nf,3]
        (same semantics as loading an OBJ holding them)."""
        V = np.ascontiguousarray(V, dtype=np.float32).reshape(-1, 3)
        F = np.ascontiguousarray(F, dtype=np.uint32).reshape(-1, 3)
        N.check(N.lib().se3cq_load_mesh(self._w, int(geo), V.ctypes.data_as(N.c_float_p), V.shape[0],
                                        F.ctypes.data_as(N.c_u32_p), F.shape[0]))
        if geo == self.GEO_ROB:
            self._state = get_identity_state_vector()

    def enforceRobotCenter(self, center):
        c = _vec3(center)
        N.check(N.lib().se3cq_enforce_robot_center(self._w, N.dptr(c)))

    def scaleToUnit(self):
        N.check(N.lib().se3cq_scale_to_unit(self._w))

    def angleModel(self, latitude, longitude):
        N.check(N.lib().se3cq_angle_model(self._w, float(latitude), float(longitude)))

    def set_perturbation(self, pert):
        q = _state(pert)
        N.check(N.lib().se3cq_set_perturbation(self._w, N.dptr(q)))

    @property
    def perturbation(self):
        q = np.zeros(7)
        N.check(N.lib().se3cq_get_perturbation(self._w, N.dptr(q)))
        return q

    @property
    def state(self):
        return self._state.copy()

    @state.setter
    def state(self, q):
        self._state = _state(q).copy()

    @property
    def recommended_cres(self):
        v = ctypes.c_double(0.0)
        N.check(N.lib().se3cq_get_recommended_cres(self._w, ctypes.byref(v)))
        return v.value

    @recommended_cres.setter
    def recommended_cres(self, v):
        N.check(N.lib().se3cq_set_recommended_cres(self._w, float(v)))

    @property
    def scene_scale(self):
        v = ctypes.c_float(0.0)
        N.check(N.lib().se3cq_get_scene_scale(self._w, ctypes.byref(v)))
        return v.value

    def _matrix(self, geo):
        m = np.zeros((4, 4))
        N.check(N.lib().se3cq_get_matrix(self._w, geo, N.dptr(m)))
        return m

    @property
    def scene_matrix(self):
        return self._matrix(0)

    @property
    def robot_matrix(self):
        return self._matrix(1)

    def get_ompl_center(self, geo=GEO_ROB):
        c = np.zeros(3, np.float32)
        N.check(N.lib().se3cq_get_ompl_center(self._w, int(geo), c.ctypes.data_as(N.c_float_p)))
        return c.astype(np.float64)

    def cd_vertices(self, geo):
        """Unit-world collision vertices (binary64), as CDModel stores them."""
        nv, nf = ctypes.c_int64(0), ctypes.c_int64(0)
        N.check(N.lib().se3cq_get_mesh_size(self._w, geo, ctypes.byref(nv), ctypes.byref(nf)))
        out = np.zeros((nv.value, 3))
        N.check(N.lib().se3cq_get_cd_vertices(self._w, geo, N.dptr(out)))
        return out

    def mesh(self, geo):
        nv, nf = ctypes.c_int64(0), ctypes.c_int64(0)
        N.check(N.lib().se3cq_get_mesh_size(self._w, geo, ctypes.byref(nv), ctypes.byref(nf)))
        V = np.zeros((nv.value, 3), np.float32)
        F = np.zeros((nf.value, 3), np.uint32)
        N.check(N.lib().se3cq_get_mesh(self._w, geo, V.ctypes.data_as(N.c_float_p), F.ctypes.data_as(N.c_u32_p)))
        return V, F

    def bvh_info(self, geo):
        d, n = ctypes.c_int32(0), ctypes.c_int64(0)
        N.check(N.lib().se3cq_get_bvh_info(self._w, geo, ctypes.byref(d), ctypes.byref(n)))
        return d.value, n.value

    # ---- validity (module.cc:138) ---------------------------------------
    def is_valid_states(self, qs):
        """Batched UnitWorld::isValid: bool array, True = collision free."""
        qs = _states(qs)
        out = np.zeros(qs.shape[0], np.uint8)
        N.check(N.lib().se3cq_valid_states_host(self._w, N.dptr(qs), qs.shape[0],
                                                out.ctypes.data_as(N.c_u8_p)))
        return out.astype(bool)

    def is_valid_state(self, state):
        return bool(self.is_valid_states(_state(state))[0])

    def contacts(self, qs, cap=1024):
        """All intersecting (env triangle, robot triangle) pairs per state with
        their intersection segments.  Returns (counts int32[n], pairs
        int32[n, cap, 2], segs float64[n, cap, 6], flags int8[n, cap])."""
        qs = _states(qs)
        n = qs.shape[0]
        counts = np.zeros(n, np.int32)
        pairs = np.zeros((n, cap, 2), np.int32)
        segs = np.zeros((n, cap, 6))
        flags = np.zeros((n, cap), np.int8)
        N.check(N.lib().se3cq_contacts_host(self._w, N.dptr(qs), n, int(cap), pairs.ctypes.data_as(N.c_i32_p),
                                            N.dptr(segs), flags.ctypes.data_as(N.c_i8_p),
                                            counts.ctypes.data_as(N.c_i32_p)))
        return counts, pairs, segs, flags

    def intersecting_segments(self, unitq):
        """UnitWorld::intersectingSegments (unit_world.cc:959-1012): segment
        start points [m, 3], end points [m, 3], lengths [m] and the
        (env face, robot face) pairs [m, 2] of the intersecting triangles."""
        cap = 1024
        while True:
            counts, pairs, segs, _ = self.contacts(_state(unitq), cap)
            m = int(counts[0])
            if m <= cap:
                break
            cap = m
        p0, p1 = segs[0, :m, :3].copy(), segs[0, :m, 3:].copy()
        mag = np.sqrt(((p1 - p0) ** 2).sum(axis=1))
        return p0, p1, mag, pairs[0, :m].astype(np.int32)

    def is_disentangled_states(self, qs):
        """Batched UnitWorld::isDisentangled: True where the two models'
        bounding boxes are disjoint (unit_world.cc:249-264)."""
        qs = _states(qs)
        out = np.zeros(qs.shape[0], np.uint8)
        N.check(N.lib().se3cq_disentangled_host(self._w, N.dptr(qs), qs.shape[0], out.ctypes.data_as(N.c_u8_p)))
        return out.astype(bool)

    def is_disentangled(self, state):
        return bool(self.is_disentangled_states(_state(state))[0])

    def transit_state(self, state, action, transit_magnitude, verify_delta):
        raise NotImplementedError("transit_state (discrete RL actions) is outside this engine's scope")

    # ---- segments (module.cc:141-156) -----------------------------------
    def transit_states_to_with_contact(self, froms, tos, verify_delta):
        """Batched transitStateToWithContact.  Returns (last_free [n,7],
        first_colliding [n,7], done bool[n], last_tau f64[n], tau f64[n],
        steps int32[n])."""
        f, t = _states(froms), _states(tos)
        if f.shape != t.shape:
            raise ValueError("froms and tos must have the same shape")
        n = f.shape[0]
        lf, st = np.zeros((n, 7)), np.zeros((n, 7))
        done = np.zeros(n, np.uint8)
        lt, ta = np.zeros(n), np.zeros(n)
        steps = np.zeros(n, np.int32)
        N.check(N.lib().se3cq_transit_host(self._w, N.dptr(f), N.dptr(t), n, float(verify_delta), N.dptr(lf),
                                           N.dptr(st), done.ctypes.data_as(N.c_u8_p), N.dptr(lt), N.dptr(ta),
                                           steps.ctypes.data_as(N.c_i32_p)))
        return lf, st, done.astype(bool), lt, ta, steps

    def transit_state_to_with_contact(self, frm, to, verify_delta):
        lf, st, done, lt, ta, _ = self.transit_states_to_with_contact(_state(frm), _state(to), verify_delta)
        return lf[0], st[0], bool(done[0]), float(np.float32(lt[0])), float(np.float32(ta[0]))

    def transit_state_to(self, frm, to, verify_delta):
        lf, _, done, lt, _ = self.transit_state_to_with_contact(frm, to, verify_delta)
        return lf, done, lt

    def is_valid_transitions(self, froms, tos, initial_verify_delta):
        f, t = _states(froms), _states(tos)
        if f.shape != t.shape:
            raise ValueError("froms and tos must have the same shape")
        out = np.zeros(f.shape[0], np.uint8)
        N.check(N.lib().se3cq_valid_transitions_host(self._w, N.dptr(f), N.dptr(t), f.shape[0],
                                                     float(initial_verify_delta), out.ctypes.data_as(N.c_u8_p)))
        return out.astype(bool)

    def is_valid_transition(self, frm, to, initial_verify_delta):
        return bool(self.is_valid_transitions(_state(frm), _state(to), initial_verify_delta)[0])

    def touch_search(self, froms, tos, verify_delta, levels=20):
        """transitStateToWithContact + `levels` bisection halvings of the
        contact bracket.  Returns a dict of arrays (done, last_tau, tau, lo,
        hi, steps)."""
        f, t = _states(froms), _states(tos)
        n = f.shape[0]
        done = np.zeros(n, np.uint8)
        lt, ta, lo, hi = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
        steps = np.zeros(n, np.int32)
        N.check(N.lib().se3cq_touch_search_host(self._w, N.dptr(f), N.dptr(t), n, float(verify_delta), int(levels),
                                                done.ctypes.data_as(N.c_u8_p), N.dptr(lt), N.dptr(ta), N.dptr(lo),
                                                N.dptr(hi), steps.ctypes.data_as(N.c_i32_p)))
        return dict(done=done.astype(bool), last_tau=lt, tau=ta, lo=lo, hi=hi, steps=steps)

    def bisect_contact(self, froms, tos, lo, hi, levels=20):
        f, t = _states(froms), _states(tos)
        n = f.shape[0]
        lo_in = np.ascontiguousarray(lo, dtype=np.float64).reshape(-1)
        hi_in = np.ascontiguousarray(hi, dtype=np.float64).reshape(-1)
        lo_out, hi_out = np.zeros(n), np.zeros(n)
        N.check(N.lib().se3cq_bisect_contact_host(self._w, N.dptr(f), N.dptr(t), n, N.dptr(lo_in), N.dptr(hi_in),
                                                  int(levels), N.dptr(lo_out), N.dptr(hi_out)))
        return lo_out, hi_out

    def transit_state_by(self, frm, tr, aa, verify_delta):
        """transitStateBy (unit_world.cc:489-506): the end state rotates the
        (unnormalised) start quaternion by AngleAxis(|aa|, aa/|aa|)."""
        f = _state(frm)
        tr, aa = _vec3(tr), _vec3(aa)
        ang = float(np.sqrt((aa[0] * aa[0] + aa[1] * aa[1]) + aa[2] * aa[2]))
        axis = aa / ang if ang > 0 else aa
        ha = 0.5 * ang
        qa = np.array([np.cos(ha), *(np.sin(ha) * axis)])
        qb = f[3:]
        w1, x1, y1, z1 = qa
        w2, x2, y2, z2 = qb
        rot = np.array([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2,
                        w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                        w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
                        w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2])
        to = np.concatenate([f[:3] + tr, rot])
        return self.transit_state_to(f, to, verify_delta)

    # ---- unit-state translation (module.cc:157-158) ---------------------
    def translate_to_unit_state(self, state):
        q = _state(state)
        out = np.zeros(7)
        N.check(N.lib().se3cq_translate_to_unit(self._w, N.dptr(q), 1, N.dptr(out)))
        return out

    def _translate(self, fn, Q, *extra):
        q = _states(Q)
        out = np.zeros_like(q)
        N.check(getattr(N.lib(), fn)(self._w, N.dptr(q), q.shape[0], *extra, N.dptr(out)))
        return out

    def translate_unit_to_ompl(self, Q, to_angle_axis=False):
        """translateUnitStateToOMPLState (unit_world.cc:558-605)."""
        return self._translate("se3cq_translate_unit_to_ompl", Q, int(bool(to_angle_axis)))

    def translate_ompl_to_unit(self, Q):
        """translateOMPLStateToUnitState (unit_world.cc:689-711)."""
        return self._translate("se3cq_translate_ompl_to_unit", Q)

    def translate_vanilla_to_ompl(self, Q):
        return self._translate("se3cq_translate_vanilla_to_ompl", Q)

    def translate_ompl_to_vanilla(self, Q):
        return self._translate("se3cq_translate_ompl_to_vanilla", Q)

    def translate_vanilla_to_unit(self, Q):
        return self._translate("se3cq_translate_vanilla_to_unit", Q)

    def translate_vanilla_pts_to_unit(self, geo, pts):
        p = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
        out = np.zeros_like(p)
        N.check(N.lib().se3cq_translate_vanilla_pts_to_unit(self._w, int(geo), N.dptr(p), p.shape[0], N.dptr(out)))
        return out

    def translate_from_unit_state(self, state):
        q = _state(state)
        out = np.zeros(7)
        N.check(N.lib().se3cq_translate_from_unit(self._w, N.dptr(q), 1, N.dptr(out)))
        return out

    # ---- visibility (module.cc:168-184) ---------------------------------
    def calculate_visibility_matrix(self, qs, is_unit_states, verify_magnitude):
        q = _states(qs)
        n = q.shape[0]
        out = np.zeros((n, n), np.int8)
        N.check(N.lib().se3cq_visibility_matrix_host(self._w, N.dptr(q), n, int(bool(is_unit_states)),
                                                     float(verify_magnitude), out.ctypes.data_as(N.c_i8_p)))
        return out

    def calculate_visibility_matrix2(self, qs0, qs0_are_unit_states, qs1, qs1_are_unit_states, verify_magnitude,
                                     enable_mt=True):
        a, b = _states(qs0), _states(qs1)
        out = np.zeros((a.shape[0], b.shape[0]), np.int8)
        N.check(N.lib().se3cq_visibility_matrix2_host(self._w, N.dptr(a), a.shape[0], int(bool(qs0_are_unit_states)),
                                                      N.dptr(b), b.shape[0], int(bool(qs1_are_unit_states)),
                                                      float(verify_magnitude), out.ctypes.data_as(N.c_i8_p)))
        return out

    def calculate_visibility_pair(self, qs0, qs0_are_unit_states, qs1, qs1_are_unit_states, verify_magnitude,
                                  enable_mt=True):
        a, b = _states(qs0), _states(qs1)
        if a.shape[0] != b.shape[0]:
            raise RuntimeError("calculate_visibility_pair needs qs0 and qs1 with the same number of rows")
        out = np.zeros((a.shape[0], 1), np.int8)
        N.check(N.lib().se3cq_visibility_pair_host(self._w, N.dptr(a), int(bool(qs0_are_unit_states)), N.dptr(b),
                                                   int(bool(qs1_are_unit_states)), a.shape[0],
                                                   float(verify_magnitude), out.ctypes.data_as(N.c_i8_p)))
        return out

    # ---- device-buffer entry points (pointers to HBM, e.g. torch .data_ptr()) --
    def disentangled_device(self, d_q, n, d_out, stream=None):
        N.check(N.lib().se3cq_disentangled_device(self._w, d_q, int(n), d_out, stream))

    def valid_states_device(self, d_q, n, d_out, stream=None):
        N.check(N.lib().se3cq_valid_states_device(self._w, d_q, int(n), d_out, stream))

    def valid_states_subset_device(self, d_q, n, d_idx, d_count, d_out, stream=None):
        """out[idx[k]] = isValid(q[idx[k]]) for k < *count (device count)."""
        N.check(N.lib().se3cq_valid_states_subset_device(self._w, d_q, int(n), d_idx, d_count, d_out, stream))

    def valid_transitions_subset_device(self, d_from, d_to, n, d_idx, d_count, initial_verify_delta, d_out,
                                        d_steps=None, stream=None):
        """out[idx[k]] = isValidTransition(from[idx[k]], to[idx[k]]) for k < *count."""
        N.check(N.lib().se3cq_valid_transitions_subset_device(self._w, d_from, d_to, int(n), d_idx, d_count,
                                                              float(initial_verify_delta), d_out, d_steps, stream))

    def transit_device(self, d_from, d_to, n, verify_delta, d_done, d_last_tau, d_tau, d_steps,
                       d_last_free=None, d_state=None, stream=None):
        """Device-pointer form of transit_states_to_with_contact (asynchronous)."""
        N.check(N.lib().se3cq_transit_device(self._w, d_from, d_to, int(n), float(verify_delta), d_last_free,
                                             d_state, d_done, d_last_tau, d_tau, d_steps, stream))

    def valid_transitions_device(self, d_from, d_to, n, initial_verify_delta, d_out, d_steps=None, stream=None):
        """Device-pointer form of is_valid_transitions (asynchronous)."""
        N.check(N.lib().se3cq_valid_transitions_device(self._w, d_from, d_to, int(n), float(initial_verify_delta),
                                                       d_out, d_steps, stream))

    def touch_search_device(self, d_from, d_to, n, verify_delta, levels, d_done, d_last_tau, d_tau, d_lo, d_hi,
                            d_steps, stream=None):
        N.check(N.lib().se3cq_touch_search_device(self._w, d_from, d_to, int(n), float(verify_delta), int(levels),
                                                  d_done, d_last_tau, d_tau, d_lo, d_hi, d_steps, stream))

    def touch_search_device_noalloc(self, d_from, d_to, n, verify_delta, levels, d_done, d_last_tau, d_tau, d_lo,
                                    d_hi, d_steps, stream=None):
        """touch_search_device with no allocation: its scratch is the library's
        static per-device array (at most SE3CQ_NOALLOC_MAX_SEGMENTS segments;
        the world must have been queried before; overlapping calls on one
        device race)."""
        N.check(N.lib().se3cq_touch_search_device_noalloc(self._w, d_from, d_to, int(n), float(verify_delta),
                                                          int(levels), d_done, d_last_tau, d_tau, d_lo, d_hi,
                                                          d_steps, stream))

    def visibility_bits_device(self, d_qs, n, verify_magnitude, d_bits, d_steps=None, stream=None):
        N.check(N.lib().se3cq_visibility_bits_device(self._w, d_qs, int(n), float(verify_magnitude), d_bits, d_steps,
                                                     stream))

    # ---- mixed batches (one persistent launch) -----------------------------------
    def batch(self, seg_from, seg_to, verify_delta, states, return_states=False):
        """Segments under transit_state_to_with_contact semantics and single
        states in one launch.  Returns a dict: done, last_tau, tau, steps (per
        segment), valid (per state), and last_free / state when asked."""
        f = _states(seg_from) if len(seg_from) else np.zeros((0, 7))
        t = _states(seg_to) if len(seg_to) else np.zeros((0, 7))
        q = _states(states) if len(states) else np.zeros((0, 7))
        if len(f) != len(t):
            raise ValueError("seg_from and seg_to must have the same number of rows")
        ns, nq = len(f), len(q)
        out = {"done": np.zeros(ns, np.uint8), "last_tau": np.zeros(ns), "tau": np.zeros(ns),
               "steps": np.zeros(ns, np.int32), "valid": np.zeros(nq, np.uint8)}
        lf = st = None
        if return_states:
            out["last_free"] = np.zeros((ns, 7))
            out["state"] = np.zeros((ns, 7))
            lf, st = N.dptr(out["last_free"]), N.dptr(out["state"])
        N.check(N.lib().se3cq_batch_host(self._w, N.dptr(f), N.dptr(t), ns, float(verify_delta), lf, st,
                                         out["done"].ctypes.data_as(N.c_u8_p), N.dptr(out["last_tau"]),
                                         N.dptr(out["tau"]), out["steps"].ctypes.data_as(N.c_i32_p), N.dptr(q), nq,
                                         out["valid"].ctypes.data_as(N.c_u8_p)))
        out["done"] = out["done"].astype(bool)
        out["valid"] = out["valid"].astype(bool)
        return out

    def batch_device(self, d_from, d_to, n_seg, verify_delta, d_done, d_last_tau, d_tau, d_steps, d_states,
                     n_states, d_valid, d_last_free=None, d_state=None, stream=None):
        """Device-pointer form of batch (asynchronous on the stream)."""
        N.check(N.lib().se3cq_batch_device(self._w, d_from, d_to, int(n_seg), float(verify_delta), d_last_free,
                                           d_state, d_done, d_last_tau, d_tau, d_steps, d_states, int(n_states),
                                           d_valid, stream))

    def batch_tune(self, d_from, d_to, n_seg, verify_delta, d_done, d_last_tau, d_tau, d_steps, d_states, n_states,
                   d_valid, d_last_free=None, d_state=None, repeats=3):
        """Search the world's leaf size, the launch shapes, the pool chunk and
        the queue depth on this batch (each candidate's outputs checked equal
        on the device) and keep the fastest for later calls.  Returns
        (config dict, ms per launch); the chosen knobs are in `tuning`."""
        cfg = N.BatchConfig()
        ms = ctypes.c_float(0.0)
        N.check(N.lib().se3cq_batch_tune(self._w, d_from, d_to, int(n_seg), float(verify_delta), d_last_free,
                                         d_state, d_done, d_last_tau, d_tau, d_steps, d_states, int(n_states),
                                         d_valid, int(repeats), ctypes.byref(cfg), ctypes.byref(ms)))
        return {k: getattr(cfg, k) for k, _ in N.BatchConfig._fields_}, ms.value

    @property
    def batch_config(self):
        cfg = N.BatchConfig()
        N.check(N.lib().se3cq_batch_get_config(self._w, ctypes.byref(cfg)))
        return {k: getattr(cfg, k) for k, _ in N.BatchConfig._fields_}

    @batch_config.setter
    def batch_config(self, c):
        cfg = N.BatchConfig(int(c.get("block_waves", 0)), int(c.get("blocks_per_cu", 0)),
                            int(c.get("states_per_lane", 1)), int(c.get("split", 0)))
        N.check(N.lib().se3cq_batch_set_config(self._w, ctypes.byref(cfg)))

    @property
    def tuning(self):
        """World-wide traversal knobs: leaf_size (triangles per BVH leaf, at
        most), pool_chunk (states per counter add of the state kernels) and
        queue_depth (segments per counter add of the segment kernels)."""
        t = N.Tuning()
        N.check(N.lib().se3cq_get_tuning(self._w, ctypes.byref(t)))
        return {k: getattr(t, k) for k, _ in N.Tuning._fields_}

    @tuning.setter
    def tuning(self, c):
        cur = self.tuning
        cur.update(c)
        t = N.Tuning(int(cur["leaf_size"]), int(cur["pool_chunk"]), int(cur["queue_depth"]))
        N.check(N.lib().se3cq_set_tuning(self._w, ctypes.byref(t)))

    def check_device(self):
        N.check(N.lib().se3cq_check(self._w))


def create_unit_world_from_arrays(env, rob):
    """A ready UnitWorld from two (V, F) meshes, following
    util._load_unit_world (src/GP/pipeline/util.py:122-128)."""
    uw = UnitWorld()
    uw.load_mesh(UnitWorld.GEO_ENV, *env)
    uw.load_mesh(UnitWorld.GEO_ROB, *rob)
    uw.scaleToUnit()
    uw.angleModel(0.0, 0.0)
    return uw


def create_unit_world(puzzle_file):
    """util.create_unit_world (src/GP/pipeline/util.py:122-140): the puzzle's
    meshes scaled to the unit world, with recommended_cres = scene_scale x the
    file's collision_resolution (0.001 when absent)."""
    from .pyse3ompl import read_puzzle
    cfg = read_puzzle(puzzle_file)
    uw = UnitWorld()
    uw.loadModelFromFile(cfg["env_fn"])
    uw.loadRobotFromFile(cfg["rob_fn"])
    uw.scaleToUnit()
    uw.angleModel(0.0, 0.0)
    cres = cfg["collision_resolution"]
    uw.recommended_cres = uw.scene_scale * (0.001 if cres is None else cres)
    return uw
