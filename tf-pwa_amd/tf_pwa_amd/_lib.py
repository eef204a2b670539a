"""ctypes binding of libtfpwa_amd.so (the C ABI declared in include/tfpwa_amd.h).

The library is built in-tree by ``tf-pwa_amd/build.py`` into ``tf-pwa_amd/lib``. There is no
fallback: if the library is missing or fails to load, every engine call raises.
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libtfpwa_amd.so")

MAX_FINAL = 3
MAX_TOPO = 3
MAX_CHAIN = 32
MAX_LS = 8
MAX_HEL = 8
MAX_J2 = 7
MAX_P = 128
MAX_LAM = 64
NF_TOPO = 18

MODE_DATA = 0
MODE_MC = 1
FLAG_GRAD = 1
FLAG_FOUT = 2
FLAG_SHAPE = 4
FLAG_F32 = 8
FLAG_F32H = 16
FLAG_F64H = 32
FUSED_MAX_P = 8  # largest model (partial waves) tpa_eval_fused accepts
CPL_SLOTS = 4  # factor slots per partial wave in TpaCouplingMap (total + up to 3 vertices)

_i32 = ctypes.c_int32


class TpaTopo(ctypes.Structure):
    _fields_ = [
        ("x", _i32), ("y", _i32), ("z", _i32),
        ("r_first", _i32),
        ("aligned", _i32 * MAX_FINAL),
    ]


class TpaChain(ctypes.Structure):
    _fields_ = [
        ("topo", _i32),
        ("top_r_first", _i32),
        ("r_y_first", _i32),
        ("jr2", _i32),
        ("n_r", _i32),
        ("hel_r", _i32 * MAX_HEL),
        ("n_ls1", _i32), ("n_ls2", _i32),
        ("l1", _i32 * MAX_LS), ("l2", _i32 * MAX_LS),
        ("act1", _i32 * MAX_LS), ("act2", _i32 * MAX_LS),
        ("k0", _i32),
        ("shape", _i32),
        ("bw_l", _i32),
        ("cg1_off", _i32), ("cg2_off", _i32),
        ("m0", ctypes.c_double), ("g0", ctypes.c_double),
        ("m_y", ctypes.c_double), ("m_z", ctypes.c_double),
        ("d", ctypes.c_double),
    ]


class TpaModel(ctypes.Structure):
    _fields_ = [
        ("n_final", _i32), ("n_topo", _i32), ("n_chain", _i32), ("P", _i32), ("lam", _i32),
        ("lam_pad", _i32),
        ("top_j2", _i32), ("n_top", _i32),
        ("hel_top", _i32 * MAX_HEL),
        ("final_j2", _i32 * MAX_FINAL),
        ("n_fh", _i32 * MAX_FINAL),
        ("hel_f", (_i32 * MAX_HEL) * MAX_FINAL),
        ("ref_topo", _i32 * MAX_FINAL),
        ("random_z", _i32),
        ("n_cg", _i32),
        ("topo", TpaTopo * MAX_TOPO),
        ("chain", TpaChain * MAX_CHAIN),
    ]


MAX4_TOPO = 8
MAX4_CHAIN = 32
NF4_TOPO = 15


class Tpa4Topo(ctypes.Structure):
    _fields_ = [
        ("core", _i32 * 3),
        ("out", (_i32 * 2) * 3),
        ("parent", _i32 * 3),
        ("parent_out", _i32 * 3),
    ]


class Tpa4Chain(ctypes.Structure):
    _fields_ = [
        ("topo", _i32),
        ("swap", _i32 * 3),
        ("out_slot", (_i32 * 2) * 3),
        ("n_ls", _i32 * 3),
        ("l", (_i32 * MAX_LS) * 3),
        ("act", (_i32 * MAX_LS) * 3),
        ("cg_off", _i32 * 3),
        ("k0", _i32),
        ("r_j2", _i32 * 2), ("r_n", _i32 * 2),
        ("r_hel", (_i32 * MAX_HEL) * 2),
        ("r_shape", _i32 * 2), ("r_bw_l", _i32 * 2),
        ("r_m0", ctypes.c_double * 2), ("r_g0", ctypes.c_double * 2),
        ("r_q0m", (ctypes.c_double * 2) * 2),
        ("d", ctypes.c_double),
    ]


class TpaModel4(ctypes.Structure):
    _fields_ = [
        ("n_final", _i32), ("n_topo", _i32), ("n_chain", _i32), ("P", _i32), ("lam", _i32),
        ("lam_pad", _i32),
        ("top_j2", _i32), ("n_top", _i32),
        ("hel_top", _i32 * MAX_HEL),
        ("random_z", _i32),
        ("n_cg", _i32),
        ("topo", Tpa4Topo * MAX4_TOPO),
        ("chain", Tpa4Chain * MAX4_CHAIN),
    ]


class TpaCouplingMap(ctypes.Structure):
    _fields_ = [
        ("P", _i32), ("n_factor", _i32), ("n_var", _i32), ("reserved", _i32),
        ("wave_factor", ctypes.POINTER(_i32)),
        ("wave_const", ctypes.POINTER(ctypes.c_double)),
        ("factor_polar", ctypes.POINTER(_i32)),
        ("factor_var", ctypes.POINTER(_i32)),
        ("factor_fixed", ctypes.POINTER(ctypes.c_double)),
        ("bound_kind", ctypes.POINTER(_i32)),
        ("bound_ab", ctypes.POINTER(ctypes.c_double)),
    ]


SHAPE_MAX_CHAIN = 32
SHAPE_MAX_LS = 8
SHAPE_MAX_P = 128


class TpaShapeMap(ctypes.Structure):
    _fields_ = [
        ("n_chain", _i32), ("P", _i32),
        ("mass_var", _i32 * SHAPE_MAX_CHAIN), ("width_var", _i32 * SHAPE_MAX_CHAIN),
        ("mass_fixed", ctypes.c_double * SHAPE_MAX_CHAIN),
        ("width_fixed", ctypes.c_double * SHAPE_MAX_CHAIN),
        ("m_top", ctypes.c_double),
        ("m_bach", ctypes.c_double * SHAPE_MAX_CHAIN), ("m_y", ctypes.c_double * SHAPE_MAX_CHAIN),
        ("m_z", ctypes.c_double * SHAPE_MAX_CHAIN), ("d", ctypes.c_double * SHAPE_MAX_CHAIN),
        ("n_l1", _i32 * SHAPE_MAX_CHAIN), ("n_l2", _i32 * SHAPE_MAX_CHAIN),
        ("l1", (_i32 * SHAPE_MAX_LS) * SHAPE_MAX_CHAIN), ("l2", (_i32 * SHAPE_MAX_LS) * SHAPE_MAX_CHAIN),
        ("act1", (_i32 * SHAPE_MAX_LS) * SHAPE_MAX_CHAIN),
        ("act2", (_i32 * SHAPE_MAX_LS) * SHAPE_MAX_CHAIN),
        ("wave_chain", _i32 * SHAPE_MAX_P), ("wave_ls1", _i32 * SHAPE_MAX_P),
        ("wave_ls2", _i32 * SHAPE_MAX_P),
    ]


class EngineError(RuntimeError):
    pass


_lib = None


def lib():
    """Load the engine library once; raise if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise EngineError(
            "libtfpwa_amd.so is not built ({}); run `python tf-pwa_amd/build.py` or "
            "`__graft_entry__.build()`".format(LIB_PATH))
    L = ctypes.CDLL(LIB_PATH)
    vp, dp, i64 = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int64
    sig = {
        "tpa_abi_version": ([], ctypes.c_int),
        "tpa_last_error": ([], ctypes.c_char_p),
        "tpa_device_count": ([ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "tpa_model_create": ([vp, i64, dp, ctypes.c_int, ctypes.POINTER(vp)], ctypes.c_int),
        "tpa_model_destroy": ([vp], ctypes.c_int),
        "tpa_model_update": ([vp, vp, i64], ctypes.c_int),
        "tpa_model4_create": ([vp, i64, dp, ctypes.c_int, ctypes.POINTER(vp)], ctypes.c_int),
        "tpa_model4_update": ([vp, vp, i64], ctypes.c_int),
        "tpa_events_create": ([vp, dp, i64, dp, ctypes.POINTER(vp)], ctypes.c_int),
        "tpa_events_destroy": ([vp], ctypes.c_int),
        "tpa_events_generate": ([vp, ctypes.c_double, dp, i64, ctypes.c_uint64, ctypes.c_uint64,
                                 dp, dp, ctypes.POINTER(vp)], ctypes.c_int),
        "tpa_events_concat": ([ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(vp)], ctypes.c_int),
        "tpa_events_select": ([vp, ctypes.POINTER(i64), i64, dp, ctypes.POINTER(vp)],
                              ctypes.c_int),
        "tpa_events_size": ([vp], i64),
        "tpa_events_set_weight": ([vp, dp], ctypes.c_int),
        "tpa_events_get_angles": ([vp, dp, i64], ctypes.c_int),
        "tpa_events_build": ([vp, ctypes.c_int], ctypes.c_int),
        "tpa_events_keep_f32": ([vp, ctypes.c_int], ctypes.c_int),
        "tpa_events_keep_half": ([vp, ctypes.c_int], ctypes.c_int),
        "tpa_events_half_residual": ([vp, dp], ctypes.c_int),
        "tpa_events_get_half": ([vp, ctypes.POINTER(ctypes.c_float), i64,
                                 ctypes.POINTER(ctypes.c_float), i64,
                                 ctypes.POINTER(ctypes.c_int), ctypes.c_int], ctypes.c_int),
        "tpa_pool_keep_f32": ([vp, ctypes.c_int], ctypes.c_int),
        "tpa_events_get_pw": ([vp, dp, i64], ctypes.c_int),
        "tpa_eval": ([vp, dp, ctypes.c_int, ctypes.c_int, dp, dp], ctypes.c_int),
        "tpa_eval_noalloc": ([vp, dp, ctypes.c_int, ctypes.c_int, vp, vp], ctypes.c_int),
        "tpa_timing_enable": ([ctypes.c_int], ctypes.c_int),
        "tpa_timing_read": ([dp, ctypes.POINTER(i64)], ctypes.c_int),
        "tpa_timing_reset": ([], ctypes.c_int),
        "tpa_device_synchronize": ([], ctypes.c_int),
        "tpa_gram": ([vp, ctypes.c_double, dp, dp], ctypes.c_int),
        "tpa_gram_f64": ([vp, dp, dp], ctypes.c_int),
        "tpa_calib_copy": ([i64, ctypes.c_int, dp], ctypes.c_int),
        "tpa_hess": ([vp, dp, ctypes.c_int, dp], ctypes.c_int),
        "tpa_hess_shape": ([vp, dp, ctypes.c_int, dp], ctypes.c_int),
        "tpa_pool_add_events": ([vp, vp, dp, ctypes.POINTER(i64)], ctypes.c_int),
        "tpa_eval_fused": ([vp, dp, ctypes.c_int, ctypes.c_int, dp], ctypes.c_int),
        "tpa_eval_recompute": ([vp, dp, ctypes.c_int, ctypes.c_int, i64, dp], ctypes.c_int),
        "tpa_fused_source": ([vp, i64, dp, ctypes.c_int, ctypes.c_char_p, i64,
                              ctypes.POINTER(i64)], ctypes.c_int),
        "tpa_fused_compile_seconds": ([vp, dp], ctypes.c_int),
        "tpa_pool_hess": ([vp, i64, i64, dp, dp], ctypes.c_int),
        "tpa_pool_create": ([vp, i64, ctypes.c_int, ctypes.POINTER(vp)], ctypes.c_int),
        "tpa_pool_destroy": ([vp], ctypes.c_int),
        "tpa_pool_add": ([vp, dp, i64, dp, ctypes.POINTER(i64)], ctypes.c_int),
        "tpa_pool_release": ([vp, i64], ctypes.c_int),
        "tpa_pool_free_events": ([vp], i64),
        "tpa_pool_eval": ([vp, ctypes.c_int, ctypes.POINTER(i64), dp, ctypes.c_int, dp, dp],
                          ctypes.c_int),
        "tpa_server_create": ([vp, ctypes.POINTER(TpaCouplingMap), dp, ctypes.c_double,
                               ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                               ctypes.POINTER(vp)], ctypes.c_int),
        "tpa_server_destroy": ([vp], ctypes.c_int),
        "tpa_fcn_create": ([ctypes.POINTER(TpaCouplingMap), ctypes.POINTER(vp), ctypes.c_int,
                            ctypes.POINTER(vp), ctypes.c_int, dp, ctypes.c_double,
                            ctypes.POINTER(vp)], ctypes.c_int),
        "tpa_fcn_destroy": ([vp], ctypes.c_int),
        "tpa_fcn_nll_grad": ([vp, dp, dp, dp], ctypes.c_int),
        "tpa_fcn_set_shape_map": ([vp, vp, ctypes.POINTER(TpaShapeMap)], ctypes.c_int),
        "tpa_fcn_cache_rebuilds": ([vp, ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
        "tpa_fcn_set_data_storage": ([vp, ctypes.c_int], ctypes.c_int),
        "tpa_fcn_set_data_half": ([vp, ctypes.c_int], ctypes.c_int),
        "tpa_fcn_set_mc_half": ([vp, ctypes.c_int], ctypes.c_int),
        "tpa_fcn_minimize": ([vp, dp, ctypes.c_double, ctypes.c_int, dp, dp,
                              ctypes.POINTER(ctypes.c_int), ctypes.POINTER(i64),
                              ctypes.POINTER(i64)], ctypes.c_int),
        "tpa_model_set_lineshapes": ([vp, ctypes.c_int, dp, dp], ctypes.c_int),
        "tpa_server_nll_grad": ([vp, i64, i64, ctypes.c_double, dp, dp, dp], ctypes.c_int),
        "tpa_server_nll_grad_hess": ([vp, i64, i64, ctypes.c_double, dp, dp, dp, dp],
                                     ctypes.c_int),
        "tpa_server_submit": ([vp, i64, i64, ctypes.c_double, dp, ctypes.POINTER(ctypes.c_int)],
                              ctypes.c_int),
        "tpa_server_poll": ([vp, ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                             ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "tpa_server_result": ([vp, ctypes.c_int, dp, dp, ctypes.POINTER(ctypes.c_int),
                               ctypes.POINTER(i64), ctypes.POINTER(i64)], ctypes.c_int),
        "tpa_server_latencies": ([vp, dp, i64, ctypes.POINTER(i64)], ctypes.c_int),
        "tpa_server_counters": ([vp, ctypes.POINTER(i64), ctypes.POINTER(i64),
                                 ctypes.POINTER(i64), dp], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    if L.tpa_abi_version() != 1:
        raise EngineError("unexpected ABI version {}".format(L.tpa_abi_version()))
    _lib = L
    return L


def check(rc):
    if rc != 0:
        msg = lib().tpa_last_error().decode(errors="replace")
        raise EngineError("engine call failed ({}): {}".format(rc, msg))


def dptr(a):
    """Pointer to a C-contiguous float64 numpy array (or None)."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def exported_symbols():
    """Names declared in include/tfpwa_amd.h that the loaded library must export."""
    return [
        "tpa_abi_version", "tpa_last_error", "tpa_device_count", "tpa_model_create",
        "tpa_model_destroy", "tpa_model_update", "tpa_model4_create", "tpa_model4_update",
        "tpa_events_create", "tpa_events_generate", "tpa_events_select", "tpa_events_concat",
        "tpa_events_destroy",
        "tpa_events_size", "tpa_events_set_weight", "tpa_events_get_angles",
        "tpa_events_build", "tpa_events_get_pw", "tpa_eval", "tpa_timing_enable",
        "tpa_timing_read", "tpa_timing_reset", "tpa_device_synchronize", "tpa_calib_copy",
        "tpa_gram", "tpa_gram_f64", "tpa_hess", "tpa_pool_hess", "tpa_pool_create", "tpa_pool_destroy", "tpa_pool_add", "tpa_pool_release",
        "tpa_pool_free_events", "tpa_pool_eval", "tpa_server_create", "tpa_server_destroy",
        "tpa_server_nll_grad", "tpa_server_nll_grad_hess", "tpa_server_submit", "tpa_server_poll", "tpa_server_result",
        "tpa_server_latencies", "tpa_server_counters", "tpa_fcn_create", "tpa_fcn_destroy",
        "tpa_fcn_nll_grad", "tpa_fcn_set_shape_map", "tpa_model_set_lineshapes", "tpa_fcn_minimize",
        "tpa_fcn_cache_rebuilds", "tpa_events_keep_f32", "tpa_fcn_set_data_storage",
        "tpa_pool_keep_f32", "tpa_eval_noalloc", "tpa_hess_shape", "tpa_eval_fused",
        "tpa_fused_source", "tpa_fused_compile_seconds", "tpa_pool_add_events",
        "tpa_events_keep_half", "tpa_events_half_residual", "tpa_fcn_set_data_half",
        "tpa_events_get_half", "tpa_fcn_set_mc_half", "tpa_eval_recompute",
    ]
