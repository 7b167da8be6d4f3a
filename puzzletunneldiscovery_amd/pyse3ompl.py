"""Drop-in for the validity queries of the reference's ``pyse3ompl`` module
(lib/pyse3ompl/pyse3ompl.cc:27-104, lib/pyse3ompl/ompldriver.h:62-120):
``OmplDriver.validate_states`` and ``OmplDriver.validate_motion_pairs``
(ompldriver.cc:566-620), the two calls the pipeline makes without planning
(src/GP/pipeline/solve2.py:54 screens predicted key configurations, 219
screens all pairs of key configurations).

The reference answers them with ompl.app's FCL checker on OMPL-convention
states (robot mesh centred at its vertex mean) and OMPL's discrete motion
validator.  Here the states are translated to the unit world on the GPU
(translateOMPLStateToUnitState) and checked by the engine's kernels; the
motion samples are OMPL's own (OMPL 1.4.1 interpolation in OMPL
coordinates, k / nd for 0 < k < nd, plus the end state).

Planning (solve, optimize, presample, graph merging) is not part of this
engine: those members raise NotImplementedError.
"""
import configparser
import ctypes
import math
import os

import numpy as np

from . import _native as N
from . import pyosr

PLANNER_RRT_CONNECT = 0
PLANNER_RRT = 1
PLANNER_BKPIECE1 = 2
PLANNER_LBKPIECE1 = 3
PLANNER_KPIECE1 = 4
PLANNER_SBL = 5
PLANNER_EST = 6
PLANNER_PRM = 7
PLANNER_BITstar = 8
PLANNER_PDST = 9
PLANNER_TRRT = 10
PLANNER_BiTRRT = 11
PLANNER_LazyRRT = 12
PLANNER_LazyLBTRRT = 13
PLANNER_SPARS = 14
PLANNER_ReRRT = 15
PLANNER_RDT = PLANNER_ReRRT
PLANNER_RDT_CONNECT = 17
PLANNER_STRIDE = 18
PLANNER_FMT = 19
VSTATE_SAMPLER_UNIFORM = 0
VSTATE_SAMPLER_GAUSSIAN = 1
VSTATE_SAMPLER_OBSTACLE = 2
VSTATE_SAMPLER_MAXCLEARANCE = 3
VSTATE_SAMPLER_PROXY = 4
MODEL_PART_ENV = 0
MODEL_PART_ROB = 1
INIT_STATE = 0
GOAL_STATE = 1
EXACT_SOLUTION = 6  # ompl::base::PlannerStatus::EXACT_SOLUTION


def _ompl_states(q):
    q = np.ascontiguousarray(q, dtype=np.float64)
    if q.ndim == 1:
        q = q.reshape(1, -1)
    if q.shape[1] != 7:
        raise ValueError("OMPL states must have 7 columns (x, y, z, qx, qy, qz, qw)")
    return q


def longest_valid_segment(mins, maxs, cdres):
    """getMaximumExtent() * cdres of OMPL's SE(3) space with these bounds:
    the bounds' diagonal plus pi / 2, times the resolution."""
    lo = np.ascontiguousarray(mins, dtype=np.float64).reshape(3)
    hi = np.ascontiguousarray(maxs, dtype=np.float64).reshape(3)
    out = ctypes.c_double(0.0)
    N.check(N.lib().se3cq_ompl_longest_valid_segment(N.dptr(lo), N.dptr(hi), float(cdres), ctypes.byref(out)))
    return out.value


class OmplDriver:
    """OmplDriver of lib/pyse3ompl (validity queries only)."""

    def __init__(self):
        self._files = ["", ""]
        self._states = {}
        self._mins = None
        self._maxs = None
        self._cdres = math.nan
        self._planner = (PLANNER_RDT, VSTATE_SAMPLER_UNIFORM, "", 1)
        self._options = []
        self._uw = None

    @classmethod
    def from_unit_world(cls, uw):
        """A driver over an already loaded world (scaleToUnit and angleModel
        done), e.g. one built from vertex arrays."""
        d = cls()
        d._uw = uw
        return d

    # ---- configuration (ompldriver.h:70-120) --------------------------------------
    def set_planner(self, planner_id, valid_state_sampler_id, sample_injection_fn, rdt_k_nearest):
        self._planner = (int(planner_id), int(valid_state_sampler_id), str(sample_injection_fn),
                         int(rdt_k_nearest))

    def set_model_file(self, part_id, fn):
        if part_id < 0 or part_id > MODEL_PART_ROB:
            raise RuntimeError("OmplDriver::setModelFile: invalid part_id " + str(part_id))
        self._files[part_id] = str(fn)
        self._uw = None

    def set_state(self, state_type, tr, rot_axis, rot_angle):
        if state_type < 0 or state_type > GOAL_STATE:
            raise RuntimeError("OmplDriver::setState: invalid state_type " + str(state_type))
        self._states[state_type] = (np.asarray(tr, np.float64).reshape(3),
                                    np.asarray(rot_axis, np.float64).reshape(3), float(rot_angle))

    def set_bb(self, mins, maxs):
        self._mins = np.asarray(mins, np.float64).reshape(3).copy()
        self._maxs = np.asarray(maxs, np.float64).reshape(3).copy()

    def set_cdres(self, cdres):
        self._cdres = float(cdres)

    def set_option_vector(self, ovec):
        self._options = list(ovec)

    # ---- the engine ------------------------------------------------------------------
    def _world(self):
        if self._uw is None:
            if not self._files[MODEL_PART_ENV] or not self._files[MODEL_PART_ROB]:
                raise RuntimeError("Failed to load rob/env gemoetry")
            uw = pyosr.UnitWorld()
            uw.loadModelFromFile(self._files[MODEL_PART_ENV])
            uw.loadRobotFromFile(self._files[MODEL_PART_ROB])
            uw.scaleToUnit()
            uw.angleModel(0.0, 0.0)
            self._uw = uw
        return self._uw

    @property
    def unit_world(self):
        return self._world()

    @property
    def longest_valid_segment(self):
        if self._mins is None:
            raise RuntimeError("OmplDriver: set_bb must be called before motions are validated")
        return longest_valid_segment(self._mins, self._maxs, self._cdres)

    # ---- validity (ompldriver.cc:566-620) ---------------------------------------------
    def validate_states(self, qs0):
        """int32 [n]: 1 where the OMPL state is collision free."""
        q = _ompl_states(qs0)
        out = np.zeros(len(q), np.int32)
        w = self._world()
        if len(q):
            N.check(N.lib().se3cq_ompl_validate_states_host(w._w, N.dptr(q), len(q),
                                                            out.ctypes.data_as(N.c_i32_p)))
        return out

    def validate_motion_pairs(self, qs0, qs1, return_samples=False):
        """int32 [min(n0, n1)]: 1 where OMPL's discrete motion validator
        accepts qs0[i] -> qs1[i] (the start state is not checked).  With
        return_samples, also the states each motion costs when valid."""
        q0 = _ompl_states(qs0)
        q1 = _ompl_states(qs1)
        n = min(len(q0), len(q1))
        q0 = np.ascontiguousarray(q0[:n])
        q1 = np.ascontiguousarray(q1[:n])
        lvs = self.longest_valid_segment
        out = np.zeros(n, np.int32)
        sam = np.zeros(n, np.int32)
        w = self._world()
        if n:
            N.check(N.lib().se3cq_ompl_validate_motions_host(w._w, N.dptr(q0), N.dptr(q1), n, lvs,
                                                             out.ctypes.data_as(N.c_i32_p),
                                                             sam.ctypes.data_as(N.c_i32_p)))
        return (out, sam) if return_samples else out

    def validate_states_device(self, d_q, n, d_out, stream=None):
        """Device-pointer form: d_q float64 [n, 7], d_out int32 [n]."""
        N.check(N.lib().se3cq_ompl_validate_states_device(self._world()._w, d_q, int(n), d_out, stream))

    def validate_motion_pairs_device(self, d_q0, d_q1, n, d_out, d_samples=None, stream=None):
        """Device-pointer form: d_out int32 [n] (1 valid, 0 invalid, -1 for a
        motion needing 2^31 or more samples), d_samples int32 [n] or None."""
        N.check(N.lib().se3cq_ompl_validate_motions_device(self._world()._w, d_q0, d_q1, int(n),
                                                           self.longest_valid_segment, d_out, d_samples, stream))

    # ---- planning: outside this engine -------------------------------------------------
    def _no_planning(self, *args, **kwargs):
        raise NotImplementedError("OMPL planning is not part of this engine; only validate_states and "
                                  "validate_motion_pairs are provided")

    solve = _no_planning
    optimize = _no_planning
    presample = _no_planning
    substitute_state = _no_planning
    add_existing_graph = _no_planning
    merge_existing_graph = _no_planning
    set_sample_set = _no_planning
    set_sample_set_edges = _no_planning
    set_sample_set_flags = _no_planning
    get_sample_set_connectivity = _no_planning
    get_compact_graph = _no_planning
    get_graph_istate_indices = _no_planning
    get_graph_gstate_indices = _no_planning


# ---- OMPL puzzle files -------------------------------------------------------------------
def _read_xyz(config, prefix):
    return np.array([config.getfloat("problem", prefix + "." + c) for c in "xyz"], np.float64)


def read_puzzle(fn):
    """An OMPL puzzle .cfg, read as src/GP/pipeline/parse_ompl.py:11-53 does:
    [problem] world and robot (paths relative to the file), start and goal
    (x, y, z, axis.x/y/z, theta), volume.min / volume.max, and
    collision_resolution when present."""
    config = configparser.ConfigParser()
    if not config.read([fn]):
        raise RuntimeError(f"cannot read puzzle file {fn}")
    d = os.path.dirname(fn)
    out = {
        "env_fn": os.path.join(d, config.get("problem", "world")),
        "rob_fn": os.path.join(d, config.get("problem", "robot")),
        "volume_min": _read_xyz(config, "volume.min"),
        "volume_max": _read_xyz(config, "volume.max"),
        "collision_resolution": config.getfloat("problem", "collision_resolution", fallback=None),
    }
    for key in ("start", "goal"):
        out[key] = (_read_xyz(config, key), _read_xyz(config, key + ".axis"),
                    config.getfloat("problem", key + ".theta"))
    return out


def create_driver(puzzle, planner_id=PLANNER_RDT, sampler_id=VSTATE_SAMPLER_UNIFORM, saminj="", rdt_k=1,
                  cdres=None, bvresize=None):
    """An OmplDriver configured from a puzzle file, as
    src/GP/pipeline/se3solver.py:35-63 configures the reference's."""
    cfg = read_puzzle(puzzle)
    drv = OmplDriver()
    drv.set_planner(planner_id, sampler_id, saminj, rdt_k)
    drv.set_model_file(MODEL_PART_ENV, cfg["env_fn"])
    drv.set_model_file(MODEL_PART_ROB, cfg["rob_fn"])
    for i, key in ((INIT_STATE, "start"), (GOAL_STATE, "goal")):
        tr, axis, theta = cfg[key]
        drv.set_state(i, tr, axis, theta)
    lo, hi = cfg["volume_min"].copy(), cfg["volume_max"].copy()
    if bvresize is not None:
        lo -= bvresize
        hi += bvresize
    drv.set_bb(lo, hi)
    if cdres is None:
        cdres = cfg["collision_resolution"]
        if cdres is None:
            raise RuntimeError(f"{puzzle}: [problem] has no collision_resolution")
    drv.set_cdres(cdres)
    return drv
