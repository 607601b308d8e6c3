"""ctypes binding of libhops_mi355x.so (include/hops_mi355x.h).

The library is built in-tree by ``mesohops_amd/build.py``.  Loading fails
loudly when it is missing: there is no CPU fallback for the propagation path.
"""
import ctypes
import os

from .errors import ArenaFull, EngineError

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhops_mi355x.so")

HOPS_PREC_FP64 = 0
HOPS_PREC_MIXED = 1
HOPS_EOM_LINEAR = 0
HOPS_EOM_NONLINEAR = 1
HOPS_EOM_NORMALIZED_NONLINEAR = 2
ABI_VERSION = 4

_i32 = ctypes.c_int32
_vp = ctypes.c_void_p

# field order must match struct hops_op in include/hops_mi355x.h
_OP_INTS = ["n_state", "n_hier", "n_mode", "n_lop", "n_act", "n_first", "ell_w", "n_pat",
            "n_mode_abs", "pat_w"]
_OP_PTRS = ["aux_ptr", "aux_mode", "aux_val",
            "up_ptr", "up_aux", "up_mode",
            "dn_ptr", "dn_aux", "dn_mode", "dn_val",
            "lr_ptr", "lr_col", "lr_val", "lc_ptr", "lc_row", "lc_val",
            "mode_lrel", "mode_lact", "mode_abs", "mode_g", "mode_w",
            "abs_g", "abs_w", "abs_lact",
            "act_lrel", "act_zrow", "am_ptr", "am_mode",
            "la_ptr", "la_row", "la_col", "la_val",
            "lq_ptr", "lq_row", "lq_col", "lq_val", "lt",
            "first_aux", "first_lact", "first_mode",
            "pat_ptr", "pat_col", "pat_h", "pc_ptr", "pc_act", "pc_kind", "pc_val",
            "ell_col", "ell_val", "k0"]


class HopsOp(ctypes.Structure):
    _fields_ = ([(n, _i32) for n in _OP_INTS] + [(n + "_d", _vp) for n in _OP_PTRS]
                + [("hd_d", _vp), ("hd_n", _i32), ("hd_pad", _i32)])


class HopsBatch(ctypes.Structure):
    _fields_ = [("B", _i32), ("precision", _i32), ("eom", _i32), ("n_t", _i32),
                ("coef_stride", _i32), ("ws_stride", _i32), ("store_stride", _i32), ("lane_shift", _i32),
                ("dt_over_hbar", ctypes.c_double), ("noise_step", _i32), ("sigma_raw", _i32),
                ("phi_d", _vp), ("y0_d", _vp), ("ya_d", _vp), ("yb_d", _vp), ("acc_d", _vp),
                ("z1_d", _vp), ("z2_d", _vp), ("zmem_d", _vp), ("ws_d", _vp), ("coef_d", _vp),
                ("psi_store_d", _vp), ("tile_d", _vp), ("n_tiles", _i32), ("tile_rows", _i32),
                ("coefT_d", _vp), ("sigT_d", _vp), ("coefT_stride", _i32), ("noise_c128", _i32), ("tile_order_d", _vp),
                ("hy_d", _vp)]


class RtSystemDesc(ctypes.Structure):
    """hops_rt_system_desc_t (include/hops_runtime.h)."""
    _fields_ = [("n_state", _i32), ("n_hmode", _i32), ("n_lop", _i32), ("pad", _i32),
                ("H", _vp), ("g", _vp), ("w", _vp), ("l_of_mode", _vp), ("l_ptr", _vp),
                ("l_row", _vp), ("l_col", _vp), ("l_val", _vp), ("lt", _vp)]


class RtTrajParam(ctypes.Structure):
    """hops_rt_traj_param_t (include/hops_runtime.h)."""
    _fields_ = [("eom", _i32), ("maxhier", _i32), ("update_step", _i32), ("early_integrator", _i32),
                ("early_steps", _i32), ("inchworm_cap", _i32), ("noise_step", _i32), ("n_t", _i32),
                ("it0", _i32), ("record", _i32), ("n_aux0", _i32), ("noise_c128", _i32),
                ("delta_a", ctypes.c_double), ("delta_s", ctypes.c_double), ("f_discard", ctypes.c_double),
                ("dt", ctypes.c_double), ("markov", _vp), ("aux0_ptr", _vp), ("aux0_mode", _vp),
                ("aux0_val", _vp), ("n_filters", _i32), ("pad2", _i32), ("filter_kind", _vp), ("filter_kmax", _vp),
                ("filter_mask", _vp)]


_lib = None


def lib():
    """Load the engine library once; raise EngineError if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise EngineError("HIP engine %s is not built; run `python -m mesohops_amd.build` "
                          "(no CPU fallback exists for the propagation path)" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    L.hops_last_error.restype = ctypes.c_char_p
    L.hops_version.restype = ctypes.c_int
    L.hops_build_super.argtypes = [ctypes.POINTER(HopsOp), ctypes.c_int, _vp]
    L.hops_propagate.argtypes = [ctypes.POINTER(HopsOp), ctypes.POINTER(HopsBatch), ctypes.c_int,
                                 ctypes.c_int, ctypes.c_int, _vp]
    L.hops_propagate_noalloc.argtypes = L.hops_propagate.argtypes
    L.hops_derivative.argtypes = [ctypes.POINTER(HopsOp), ctypes.POINTER(HopsBatch), _vp, _vp, ctypes.c_int, _vp]
    L.hops_permute_c128.argtypes = [_vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_double, _vp]
    L.hops_apply_state_op.argtypes = [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp]
    L.hops_noise_embedding.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp]
    L.hops_noise_filter.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp, _vp]
    L.hops_noise_filter_c128.argtypes = L.hops_noise_filter.argtypes
    L.hops_noise_spline.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_int, _vp, _vp, _vp, _vp]
    L.hops_noise_window_mean.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp]
    L.hops_profile_enable.argtypes = [ctypes.c_int]
    L.hops_set_option.argtypes = [ctypes.c_char_p, ctypes.c_int]
    L.hops_profile_read.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_longlong)]
    _ll = ctypes.c_longlong
    for name in ("hops_slots_step",):
        getattr(L, name).argtypes = [_vp, _vp, _vp, ctypes.c_int, _vp, ctypes.c_int, _vp, ctypes.c_int,
                                     ctypes.c_int, _vp]
    L.hops_slots_pn.argtypes = [_vp, _vp, _vp, ctypes.c_int, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _vp]
    L.hops_slots_errors.argtypes = [_vp, _vp, _vp, ctypes.c_int, _vp, ctypes.c_int, _vp, ctypes.c_int,
                                    ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_double, _vp]
    L.hops_slots_build.argtypes = [_vp, _vp, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int, _vp]
    L.hops_slots_permute.argtypes = [_vp, _vp, ctypes.c_int, _vp]
    L.hops_copy_segments.argtypes = [_vp, ctypes.c_int, _vp]
    L.hops_rt_slot_noise.argtypes = [_vp, ctypes.c_int, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int]
    L.hops_rt_slot_static_basis.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, _vp, _vp, _vp]
    L.hops_rt_slot_stage_noise.argtypes = [_vp, ctypes.c_int, _vp, _vp]
    L.hops_gather_u64.argtypes = [_vp, _vp, ctypes.c_int, _vp]
    L.hops_rt_system_create.restype = _vp
    L.hops_rt_system_create.argtypes = [ctypes.POINTER(RtSystemDesc)]
    L.hops_rt_system_destroy.argtypes = [_vp]
    L.hops_rt_batch_create.restype = _vp
    L.hops_rt_batch_create.argtypes = [ctypes.c_int, _vp, ctypes.c_size_t, ctypes.c_int, ctypes.c_int, _vp]
    L.hops_rt_batch_destroy.argtypes = [_vp]
    L.hops_rt_batch_add.argtypes = [_vp, _vp, ctypes.POINTER(RtTrajParam), _vp, _vp, _vp]
    L.hops_rt_batch_remove.argtypes = [_vp, ctypes.c_int]
    L.hops_rt_slot_estimate.restype = _ll
    L.hops_rt_slot_estimate.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int]
    L.hops_rt_arena_free.restype = _ll
    L.hops_rt_arena_free.argtypes = [_vp]
    L.hops_rt_batch_timers.argtypes = [_vp, _vp]
    L.hops_rt_batch_option.argtypes = [_vp, ctypes.c_char_p, _ll]
    L.hops_rt_batch_step.argtypes = [_vp, ctypes.c_int, _vp, ctypes.c_int]
    L.hops_rt_slot_info.argtypes = [_vp, ctypes.c_int, _vp]
    L.hops_rt_slot_states.argtypes = [_vp, ctypes.c_int, _vp]
    L.hops_rt_slot_aux.argtypes = [_vp, ctypes.c_int, _vp, _vp, _vp]
    L.hops_rt_slot_phi.argtypes = [_vp, ctypes.c_int, _vp]
    L.hops_rt_slot_zmem.argtypes = [_vp, ctypes.c_int, _vp]
    L.hops_rt_slot_record.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp]
    L.hops_rt_slot_clear_records.argtypes = [_vp, ctypes.c_int]
    L.hops_rt_debug_basis.restype = _vp
    L.hops_rt_debug_basis.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, _vp, _vp, _vp]
    L.hops_rt_debug_table.restype = _ll
    L.hops_rt_debug_table.argtypes = [_vp, ctypes.c_char_p, _vp, _ll]
    L.hops_rt_debug_select.argtypes = [_vp, ctypes.POINTER(RtTrajParam), _vp, _vp, ctypes.c_double, _vp]
    L.hops_rt_debug_select_states.argtypes = [_vp, _vp]
    L.hops_rt_debug_noise.argtypes = [_vp, _vp, _vp, _vp]
    L.hops_rt_debug_set_zobrist.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_ulonglong, ctypes.c_ulonglong]
    L.hops_rt_debug_errors.restype = _ll
    L.hops_rt_debug_errors.argtypes = [_vp, ctypes.c_int, _vp, _ll]
    L.hops_rt_debug_result.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _ll]
    L.hops_rt_debug_free.argtypes = [_vp]
    L.hops_rt_debug_threshold.restype = ctypes.c_double
    L.hops_rt_debug_threshold.argtypes = [_vp, ctypes.c_longlong, ctypes.c_double, ctypes.c_double, _vp]
    if L.hops_version() != ABI_VERSION:
        raise EngineError("libhops_mi355x.so ABI %d != %d; rebuild" % (L.hops_version(), ABI_VERSION))
    _lib = L
    return L


HOPS_RT_ARENA_FULL = -2


def check(rc):
    if rc != 0:
        msg = lib().hops_last_error().decode()
        if rc == HOPS_RT_ARENA_FULL:
            raise ArenaFull(msg)
        raise EngineError(msg)


EXPORTED = ["hops_last_error", "hops_version", "hops_profile_enable", "hops_profile_read", "hops_set_option",
            "hops_build_super", "hops_propagate", "hops_propagate_noalloc", "hops_derivative", "hops_permute_c128", "hops_apply_state_op",
            "hops_noise_embedding", "hops_noise_filter", "hops_noise_filter_c128", "hops_noise_spline", "hops_noise_window_mean",
            "hops_slots_step", "hops_slots_pn", "hops_slots_errors", "hops_slots_build", "hops_slots_permute", "hops_copy_segments",
            "hops_gather_u64",
            "hops_rt_system_create", "hops_rt_system_destroy", "hops_rt_batch_create", "hops_rt_batch_destroy",
            "hops_rt_batch_add", "hops_rt_batch_remove", "hops_rt_slot_noise",
            "hops_rt_slot_static_basis", "hops_rt_slot_stage_noise", "hops_rt_slot_estimate", "hops_rt_arena_free", "hops_rt_batch_timers",
            "hops_rt_batch_option", "hops_rt_batch_step", "hops_rt_slot_info", "hops_rt_slot_states", "hops_rt_slot_aux",
            "hops_rt_slot_phi", "hops_rt_slot_zmem", "hops_rt_slot_record", "hops_rt_slot_clear_records",
            "hops_rt_debug_basis", "hops_rt_debug_table", "hops_rt_debug_select", "hops_rt_debug_select_states",
            "hops_rt_debug_result", "hops_rt_debug_free", "hops_rt_debug_threshold", "hops_rt_debug_noise",
            "hops_rt_debug_errors", "hops_rt_debug_set_zobrist"]


def set_option(name, value):
    check(lib().hops_set_option(name.encode(), int(value)))


def profile_enable(on=True):
    check(lib().hops_profile_enable(1 if on else 0))


def profile_read():
    """(ms, launches) for [prologue, stage, other, dense system block] since
    profile_enable."""
    ms = (ctypes.c_double * 4)()
    n = (ctypes.c_longlong * 4)()
    check(lib().hops_profile_read(ms, n))
    return list(ms), list(n)
