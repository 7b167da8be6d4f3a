"""ctypes binding of libse3cq.so (the gfx950 engine behind include/se3cq.h).

There is no CPU fallback: if the library is missing or fails to load, every
call raises.  Build it with ``python -c "import __graft_entry__ as g; g.build()"``
from the repository root.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libse3cq.so")

c_double_p = ctypes.POINTER(ctypes.c_double)
c_float_p = ctypes.POINTER(ctypes.c_float)
c_u32_p = ctypes.POINTER(ctypes.c_uint32)
c_u8_p = ctypes.POINTER(ctypes.c_uint8)
c_i8_p = ctypes.POINTER(ctypes.c_int8)
c_i32_p = ctypes.POINTER(ctypes.c_int32)
c_i64_p = ctypes.POINTER(ctypes.c_int64)
vp = ctypes.c_void_p


class BatchConfig(ctypes.Structure):
    """se3cq_batch_config."""
    _fields_ = [("block_waves", ctypes.c_int32), ("blocks_per_cu", ctypes.c_int32),
                ("states_per_lane", ctypes.c_int32), ("split", ctypes.c_int32)]


class Tuning(ctypes.Structure):
    """se3cq_tuning."""
    _fields_ = [("leaf_size", ctypes.c_int32), ("pool_chunk", ctypes.c_int32), ("queue_depth", ctypes.c_int32)]
i64 = ctypes.c_int64
cint = ctypes.c_int
cdbl = ctypes.c_double

# name -> argtypes (all return int status unless listed in _RESTYPE)
SIGNATURES = {
    "se3cq_last_error": [],
    "se3cq_version": [],
    "se3cq_world_create": [ctypes.POINTER(vp)],
    "se3cq_world_destroy": [vp],
    "se3cq_world_copy_from": [vp, vp],
    "se3cq_load_file": [vp, cint, ctypes.c_char_p],
    "se3cq_load_mesh": [vp, cint, c_float_p, i64, c_u32_p, i64],
    "se3cq_get_center": [vp, cint, c_float_p],
    "se3cq_get_ompl_center": [vp, cint, c_float_p],
    "se3cq_enforce_robot_center": [vp, c_double_p],
    "se3cq_scale_to_unit": [vp],
    "se3cq_angle_model": [vp, ctypes.c_float, ctypes.c_float],
    "se3cq_get_scene_scale": [vp, c_float_p],
    "se3cq_get_matrix": [vp, cint, c_double_p],
    "se3cq_get_mesh_size": [vp, cint, c_i64_p, c_i64_p],
    "se3cq_get_mesh": [vp, cint, c_float_p, c_u32_p],
    "se3cq_get_cd_vertices": [vp, cint, c_double_p],
    "se3cq_get_bvh_info": [vp, cint, c_i32_p, c_i64_p],
    "se3cq_set_perturbation": [vp, c_double_p],
    "se3cq_get_perturbation": [vp, c_double_p],
    "se3cq_set_recommended_cres": [vp, cdbl],
    "se3cq_get_recommended_cres": [vp, c_double_p],
    "se3cq_valid_states_host": [vp, c_double_p, i64, c_u8_p],
    "se3cq_valid_states_device": [vp, vp, i64, vp, vp],
    "se3cq_transit_host": [vp, c_double_p, c_double_p, i64, cdbl, c_double_p, c_double_p, c_u8_p,
                           c_double_p, c_double_p, c_i32_p],
    "se3cq_transit_device": [vp, vp, vp, i64, cdbl, vp, vp, vp, vp, vp, vp, vp],
    "se3cq_valid_transitions_host": [vp, c_double_p, c_double_p, i64, cdbl, c_u8_p],
    "se3cq_valid_transitions_device": [vp, vp, vp, i64, cdbl, vp, vp, vp],
    "se3cq_touch_search_host": [vp, c_double_p, c_double_p, i64, cdbl, cint, c_u8_p, c_double_p,
                                c_double_p, c_double_p, c_double_p, c_i32_p],
    "se3cq_touch_search_device": [vp, vp, vp, i64, cdbl, cint, vp, vp, vp, vp, vp, vp, vp],
    "se3cq_touch_search_device_noalloc": [vp, vp, vp, i64, cdbl, cint, vp, vp, vp, vp, vp, vp, vp],
    "se3cq_valid_states_subset_device": [vp, vp, i64, vp, vp, vp, vp],
    "se3cq_valid_transitions_subset_device": [vp, vp, vp, i64, vp, vp, cdbl, vp, vp, vp],
    "se3cq_bisect_contact_host": [vp, c_double_p, c_double_p, i64, c_double_p, c_double_p, cint,
                                  c_double_p, c_double_p],
    "se3cq_visibility_matrix_host": [vp, c_double_p, i64, cint, cdbl, c_i8_p],
    "se3cq_visibility_matrix2_host": [vp, c_double_p, i64, cint, c_double_p, i64, cint, cdbl, c_i8_p],
    "se3cq_visibility_pair_host": [vp, c_double_p, cint, c_double_p, cint, i64, cdbl, c_i8_p],
    "se3cq_visibility_bits_device": [vp, vp, i64, cdbl, vp, vp, vp],
    "se3cq_contacts_host": [vp, c_double_p, i64, ctypes.c_int32, c_i32_p, c_double_p, c_i8_p, c_i32_p],
    "se3cq_disentangled_host": [vp, c_double_p, i64, c_u8_p],
    "se3cq_disentangled_device": [vp, vp, i64, vp, vp],
    "se3cq_check": [vp],
    "se3cq_translate_to_unit": [vp, c_double_p, i64, c_double_p],
    "se3cq_translate_from_unit": [vp, c_double_p, i64, c_double_p],
    "se3cq_translate_ompl_to_unit": [vp, c_double_p, i64, c_double_p],
    "se3cq_translate_unit_to_ompl": [vp, c_double_p, i64, cint, c_double_p],
    "se3cq_translate_vanilla_to_ompl": [vp, c_double_p, i64, c_double_p],
    "se3cq_translate_ompl_to_vanilla": [vp, c_double_p, i64, c_double_p],
    "se3cq_translate_vanilla_to_unit": [vp, c_double_p, i64, c_double_p],
    "se3cq_translate_vanilla_pts_to_unit": [vp, cint, c_double_p, i64, c_double_p],
    "se3cq_batch_device": [vp, vp, vp, ctypes.c_int64, ctypes.c_double, vp, vp, vp, vp, vp, vp, vp, ctypes.c_int64,
                           vp, vp],
    "se3cq_batch_host": [vp, c_double_p, c_double_p, ctypes.c_int64, ctypes.c_double, c_double_p, c_double_p, c_u8_p,
                         c_double_p, c_double_p, c_i32_p, c_double_p, ctypes.c_int64, c_u8_p],
    "se3cq_batch_tune": [vp, vp, vp, ctypes.c_int64, ctypes.c_double, vp, vp, vp, vp, vp, vp, vp, ctypes.c_int64, vp,
                         ctypes.c_int32, ctypes.POINTER(BatchConfig), ctypes.POINTER(ctypes.c_float)],
    "se3cq_batch_get_config": [vp, ctypes.POINTER(BatchConfig)],
    "se3cq_batch_set_config": [vp, ctypes.POINTER(BatchConfig)],
    "se3cq_get_tuning": [vp, ctypes.POINTER(Tuning)],
    "se3cq_set_tuning": [vp, ctypes.POINTER(Tuning)],
    "se3cq_ompl_longest_valid_segment": [c_double_p, c_double_p, ctypes.c_double, c_double_p],
    "se3cq_ompl_validate_states_host": [vp, c_double_p, i64, c_i32_p],
    "se3cq_ompl_validate_states_device": [vp, vp, i64, vp, vp],
    "se3cq_ompl_validate_motions_host": [vp, c_double_p, c_double_p, i64, ctypes.c_double, c_i32_p, c_i32_p],
    "se3cq_ompl_validate_motions_device": [vp, vp, vp, i64, ctypes.c_double, vp, vp, vp],
    "se3cq_distance": [c_double_p, c_double_p, c_double_p],
    "se3cq_multi_distance": [c_double_p, c_double_p, i64, c_double_p],
    "se3cq_interpolate": [c_double_p, c_double_p, cdbl, c_double_p],
    "se3cq_apply": [c_double_p, c_double_p, c_double_p, c_double_p],
    "se3cq_apply_batch": [c_double_p, c_double_p, c_double_p, i64, c_double_p],
    "se3cq_decompose_2": [c_double_p, c_double_p, c_double_p],
    "se3cq_decompose_3": [c_double_p, c_double_p, c_double_p, c_double_p],
    "se3cq_compose_from_angleaxis": [c_double_p, cdbl, c_double_p, c_double_p],
}
# include/se3mlp.h: the learned prescreen
MLP_SIGNATURES = {
    "se3mlp_last_error": [],
    "se3mlp_create": [i64, ctypes.c_uint64, ctypes.POINTER(vp)],
    "se3mlp_destroy": [vp],
    "se3mlp_param_layout": [c_i64_p],
    "se3mlp_get_params": [vp, cint, c_float_p, c_float_p],
    "se3mlp_set_params": [vp, cint, c_float_p, c_float_p],
    "se3mlp_forward_device": [vp, vp, i64, vp, vp, vp, vp],
    "se3mlp_screen_states_device": [vp, vp, i64, ctypes.c_float, vp, vp, vp, vp],
    "se3mlp_screen_segments_device": [vp, vp, vp, i64, cint, ctypes.c_float, vp, vp, vp],
    "se3mlp_train_step_device": [vp, vp, vp, i64, ctypes.c_float, c_float_p, vp],
    "se3mlp_train_epoch_device": [vp, vp, vp, vp, i64, i64, ctypes.c_float, ctypes.c_int, vp, vp],
    "se3mlp_gemm_device": [vp, i64, vp, i64, cint, cint, cint, vp, i64, vp],
    "se3mlp_features_device": [vp, i64, vp, vp],
}
_RESTYPE = {"se3cq_last_error": ctypes.c_char_p, "se3cq_version": ctypes.c_char_p,
            "se3mlp_last_error": ctypes.c_char_p}

_lib = None


def lib():
    """Load libse3cq.so once; raise if it is absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"libse3cq.so not found at {LIB_PATH}; build it first "
                "(__graft_entry__.build()). There is no CPU fallback.")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in list(SIGNATURES.items()) + list(MLP_SIGNATURES.items()):
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _RESTYPE.get(name, ctypes.c_int)
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = lib().se3cq_last_error()
        raise RuntimeError((msg.decode() if msg else "") or f"se3cq error {rc}")
    return rc


def check_mlp(rc):
    if rc != 0:
        msg = lib().se3mlp_last_error()
        raise RuntimeError((msg.decode() if msg else "") or f"se3mlp error {rc}")
    return rc


def dptr(a):
    return a.ctypes.data_as(c_double_p)
