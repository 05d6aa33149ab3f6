"""ctypes binding of the C ABI in include/xped_hip.h and include/xped_tensor.h.

The shared library is built in-tree (``python -m xped_amd.build``).  Loading fails loudly when it
is missing: there is no CPU fallback for any data operation.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libxped_hip.so"

XH_OK, XH_EINVAL, XH_ENOMEM, XH_EHIP, XH_ENOTCONV, XH_ENODEV = 0, 1, 2, 3, 4, 5
XH_FP64, XH_FP32, XH_BF16X3, XH_OZAKI = 0, 1, 2, 3
XH_MAXLEGS = 20


class XpedError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class XpedNoDevice(XpedError):
    pass


class xh_gemm_item(C.Structure):
    _fields_ = [("a_off", C.c_int64), ("b_off", C.c_int64), ("c_off", C.c_int64), ("m", C.c_int32), ("n", C.c_int32), ("k", C.c_int32)]


class xh_factor_item(C.Structure):
    _fields_ = [("a_off", C.c_int64), ("u_off", C.c_int64), ("s_off", C.c_int64), ("vt_off", C.c_int64), ("m", C.c_int32), ("n", C.c_int32)]


class xh_permute_task(C.Structure):
    _fields_ = [("dst_off", C.c_int64), ("dst_ld", C.c_int64), ("rows", C.c_int64), ("cols", C.c_int64),
                ("nsrc", C.c_int32), ("src_begin", C.c_int32), ("ext", C.c_uint32 * XH_MAXLEGS),
                ("sstride", C.c_int64 * XH_MAXLEGS), ("is_col", C.c_uint8 * XH_MAXLEGS)]


class xh_permute_source(C.Structure):
    _fields_ = [("src_off", C.c_int64), ("src_ld", C.c_int64), ("coeff", C.c_double)]


class xh_permute_desc(C.Structure):
    _fields_ = [("nlegs", C.c_int32), ("new_rank", C.c_int32), ("ntasks", C.c_int64), ("nsources", C.c_int64),
                ("tasks", C.POINTER(xh_permute_task)), ("sources", C.POINTER(xh_permute_source))]


_P = C.c_void_p
_PP = C.POINTER(C.c_void_p)
_IP = C.POINTER(C.c_int)
_DP = C.POINTER(C.c_double)
_SZ = C.c_size_t
_SZP = C.POINTER(C.c_size_t)
_I64P = C.POINTER(C.c_int64)

# (name, restype, argtypes)
_SIGS = [
    # device layer
    ("xh_create", C.c_int, [C.c_int, _PP]),
    ("xh_destroy", C.c_int, [_P]),
    ("xh_last_error", C.c_char_p, [_P]),
    ("xh_device_count", C.c_int, [_IP]),
    ("xh_alloc", C.c_int, [_P, _SZ, _PP]),
    ("xh_free", C.c_int, [_P, _P]),
    ("xh_set_device_cache", C.c_int, [_P, _SZ]),
    ("xh_fill_ranges", C.c_int, [_P, _P, _SZ, _P, C.c_int64, C.c_double]),
    ("xh_device_cache_stats", C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(_SZ)]),
    ("xh_upload", C.c_int, [_P, _P, _P, _SZ]),
    ("xh_download", C.c_int, [_P, _P, _P, _SZ]),
    ("xh_copy", C.c_int, [_P, _P, _P, _SZ]),
    ("xh_synchronize", C.c_int, [_P]),
    ("xh_stream", _P, [_P]),
    ("xh_fill", C.c_int, [_P, _P, _SZ, C.c_double]),
    ("xh_fill_gaussian", C.c_int, [_P, _P, _SZ, C.c_uint64]),
    ("xh_range_sumsq", C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.c_int64, _P, _DP]),
    ("xh_permute_plan_create", C.c_int, [_P, C.POINTER(xh_permute_desc), _PP]),
    ("xh_permute_plan_destroy", C.c_int, [_P, _P]),
    ("xh_permute", C.c_int, [_P, _P, _P, _P]),
    ("xh_permute_plan_bytes", C.c_int, [_P, _DP]),
    ("xh_gemm_plan_create", C.c_int, [_P, C.POINTER(xh_gemm_item), C.c_int64, C.c_int, C.c_int, _PP]),
    ("xh_gemm_plan_destroy", C.c_int, [_P, _P]),
    ("xh_gemm", C.c_int, [_P, _P, _P, _P, _P, C.c_double, C.c_double, C.c_int]),
    ("xh_gemm_plan_flops", C.c_int, [_P, _DP]),
    ("xh_gemm_plan_workspace", C.c_int, [_P, C.POINTER(C.c_int64)]),
    ("xh_gemm_noalloc", C.c_int, [_P, _P, _P, _P, _P, C.c_double, C.c_double, C.c_int]),
    ("xh_gemm_gather_eligible", C.c_int, [_P, C.POINTER(C.c_int)]),
    ("xh_gather_plan_create", C.c_int, [_P, _P, _P, C.POINTER(_P)]),
    ("xh_gather_plan_destroy", C.c_int, [_P, _P]),
    ("xh_gemm_gather", C.c_int, [_P, _P, _P, _P, C.c_int64, _P, _P, C.c_double, C.c_double]),
    ("xh_scatter_plan_create", C.c_int, [_P, _P, _P, C.POINTER(_P)]),
    ("xh_scatter_plan_destroy", C.c_int, [_P, _P]),
    ("xh_gemm_gather_scatter", C.c_int, [_P, _P, _P, _P, _P, C.c_int64, _P, _P, C.c_int64, C.c_double]),
    ("xh_block_svd", C.c_int, [_P, C.POINTER(xh_factor_item), C.c_int64, _P, _P, _P, _P]),
    ("xh_block_qr", C.c_int, [_P, C.POINTER(xh_factor_item), C.c_int64, _P, _P, _P]),
    ("xh_block_eigh", C.c_int, [_P, C.POINTER(xh_factor_item), C.c_int64, _P, _P, _P]),
    ("xh_block_chol_inv", C.c_int, [_P, C.POINTER(xh_factor_item), C.c_int64, _P, _P, _DP]),
    ("xh_scale", C.c_int, [_P, _P, _SZ, C.c_double]),
    ("xh_axpby", C.c_int, [_P, _P, _P, _SZ, C.c_double, C.c_double]),
    ("xh_max_abs", C.c_int, [_P, _P, _SZ, _DP]),
    ("xh_normalize_max_abs", C.c_int, [_P, _P, _P, _SZ]),
    ("xh_normalize_status", C.c_int, [_P, C.POINTER(C.c_int64), _DP]),
    ("xh_dot", C.c_int, [_P, _P, _P, _SZ, _DP]),
    ("xh_mlp_create", C.c_int, [_P, C.c_int, C.c_int, C.c_uint64, _PP]),
    ("xh_mlp_destroy", C.c_int, [_P, _P]),
    ("xh_mlp_num_params", C.c_int, [_P, C.POINTER(C.c_int64)]),
    ("xh_mlp_get_params", C.c_int, [_P, _P, C.POINTER(C.c_float)]),
    ("xh_mlp_set_params", C.c_int, [_P, _P, C.POINTER(C.c_float)]),
    ("xh_mlp_forward", C.c_int, [_P, _P, _P, C.c_int64, _P]),
    ("xh_mlp_grad", C.c_int, [_P, _P, _P, _P, C.c_int64, _P, C.POINTER(C.c_float)]),
    ("xh_mlp_adam_step", C.c_int, [_P, _P, _P, C.c_float, C.c_int64]),
    ("xh_add_at", C.c_int, [_P, _P, C.c_int64, C.c_double]),
    ("xh_gemm_timing", C.c_int, [_P, C.c_int, C.c_int, _DP, _DP, C.POINTER(C.c_int64)]),
    ("xh_trace_marker", C.c_int, [_P, C.c_int]),
    ("xh_gemm_launch_log", C.c_int, [_P, C.c_int64, _DP, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_int64, C.POINTER(C.c_int64)]),
    ("xh_set_gemm_tag", C.c_int, [_P, C.c_int]),
    ("xh_permute_timing", C.c_int, [_P, C.c_int, C.c_int, _DP, _DP, C.POINTER(C.c_int64)]),
    ("xh_set_gemm_tiles", C.c_int, [_P, C.c_int]),
    ("xh_event_record", C.c_int, [_P, C.c_int]),
    ("xh_event_elapsed_ms", C.c_int, [_P, C.c_int, C.c_int, C.POINTER(C.c_float)]),
    ("xh_set_gemm_autotune", C.c_int, [_P, C.c_int]),
    ("xh_gemm_autotune_stats", C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    ("xh_graph_begin", C.c_int, [_P]),
    ("xh_graph_end", C.c_int, [_P, _PP]),
    ("xh_graph_launch", C.c_int, [_P, _P]),
    ("xh_graph_destroy", C.c_int, [_P, _P]),
    ("xh_graph_capturing", C.c_int, [_P]),
    # tensor layer
    ("xt_rt_create", C.c_int, [C.c_int, _PP]),
    ("xt_rt_destroy", C.c_int, [_P]),
    ("xt_last_error", C.c_char_p, [_P]),
    ("xt_rt_ctx", _P, [_P]),
    ("xt_rt_counters", C.c_int, [_P, _DP, _DP, C.POINTER(C.c_long), C.POINTER(C.c_long)]),
    ("xt_rt_clear_caches", C.c_int, [_P]),
    ("xt_rt_set_max_plans", C.c_int, [_P, C.c_int64]),
    ("xt_rt_set_fuse_permute", C.c_int, [_P, C.c_int, C.POINTER(C.c_long)]),
    ("xt_rt_cache_sizes", C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    ("xt_rt_host_stats", C.c_int, [_P, _DP, _DP, C.POINTER(C.c_long), C.POINTER(C.c_long)]),
    ("xt_rt_host_timers", C.c_int, [_P, _DP, C.c_int]),
    ("xt_rt_gemm_timing", C.c_int, [_P, C.c_int, C.c_int, _DP, _DP, C.POINTER(C.c_long)]),
    ("xt_rt_gemm_log", C.c_int, [_P, C.POINTER(C.c_int32), C.c_long]),
    ("xt_rt_marker", C.c_int, [_P, C.c_int]),
    ("xt_basis_create", C.c_int, [C.c_int, C.c_int, _PP]),
    ("xt_basis_destroy", C.c_int, [_P]),
    ("xt_basis_copy", C.c_int, [_P, _PP]),
    ("xt_basis_push_back", C.c_int, [_P, _IP, _SZ]),
    ("xt_basis_conj", C.c_int, [_P, _PP]),
    ("xt_basis_combine", C.c_int, [_P, _P, _PP]),
    ("xt_basis_forget_history", C.c_int, [_P, _PP]),
    ("xt_basis_sort", C.c_int, [_P]),
    ("xt_basis_sym", C.c_int, [_P, _IP, _IP]),
    ("xt_basis_info", C.c_int, [_P, _IP, _IP, _SZP, _SZP, _IP]),
    ("xt_basis_entry", C.c_int, [_P, C.c_int, _IP, _SZP, _SZP, _IP]),
    ("xt_basis_tree", C.c_int, [_P, C.c_int, C.c_int, _IP, _IP, _IP, _IP, C.POINTER(C.c_uint32), _IP, C.POINTER(C.c_uint64)]),
    ("xt_tree_permute", C.c_int, [_P, C.c_int, C.c_int, _IP, C.c_int, C.c_int, _IP, _IP, _IP, _IP, _DP]),
    ("xt_tree_swap", C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, _IP, _IP, _IP, _IP, _DP]),
    ("xt_wigner_3j", C.c_double, [C.c_int] * 6),
    ("xt_wigner_6j", C.c_double, [C.c_int] * 6),
    ("xt_wigner_9j", C.c_double, [C.c_int] * 9),
    ("xt_coeff", C.c_double, [C.c_int, C.c_int, C.c_int, _IP, _IP, _IP, _IP, _IP, _IP]),
    ("xt_tensor_create", C.c_int, [_P, C.c_int, C.c_int, C.POINTER(C.c_void_p), _PP]),
    ("xt_tensor_destroy", C.c_int, [_P]),
    ("xt_tensor_copy", C.c_int, [_P, _PP]),
    ("xt_tensor_shape", C.c_int, [_P, _IP, _IP, _IP, _SZP]),
    ("xt_tensor_sector", C.c_int, [_P, C.c_int, _IP, _SZP, _SZP, _SZP]),
    ("xt_tensor_leg", C.c_int, [_P, C.c_int, _PP]),
    ("xt_tensor_coupled", C.c_int, [_P, C.c_int, _PP]),
    ("xt_tensor_upload", C.c_int, [_P, _DP, _SZ]),
    ("xt_tensor_download", C.c_int, [_P, _DP, _SZ]),
    ("xt_tensor_device_ptr", _P, [_P]),
    ("xt_tensor_set_random", C.c_int, [_P, C.c_uint32, C.c_int]),
    ("xt_tensor_save", C.c_int, [_P, C.c_char_p]),
    ("xt_tensor_load", C.c_int, [_P, C.c_char_p, _PP]),
    ("xt_tensor_serialize", C.c_int, [_P, _P, _SZ, _SZP]),
    ("xt_tensor_deserialize", C.c_int, [_P, _P, _SZ, _PP]),
    ("xt_permute", C.c_int, [_P, C.c_int, _IP, C.c_int, _PP]),
    ("xt_permute_plan_stats", C.c_int, [_P, C.c_int, _IP, C.c_int, _I64P, _I64P, _I64P]),
    ("xt_permute_plan_dump", C.c_int, [_P, C.c_int, _IP, C.c_int, _I64P, _I64P, _I64P, C.POINTER(C.c_int32), _I64P, _DP]),
    ("xt_mul", C.c_int, [_P, _P, C.c_int, _PP]),
    ("xt_mul_into", C.c_int, [_P, _P, _P, C.c_int]),
    ("xt_contract", C.c_int, [_P, _IP, C.c_int, _P, _IP, C.c_int, C.c_int, C.c_int, _PP]),
    ("xt_get_permutations", C.c_int, [_IP, C.c_int, C.c_int, _IP, C.c_int, C.c_int, C.c_int, _IP, _IP, _IP, _IP, _IP, _IP]),
    ("xt_tsvd", C.c_int, [_P, _SZ, C.c_double, C.c_int, _PP, _PP, _PP, _DP, _DP]),
    ("xt_tsvd_randomized", C.c_int, [_P, _SZ, C.c_double, _IP, _IP, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_uint64, _PP, _PP, _PP,
                                     _DP, _IP, _DP]),
    ("xt_svd_full", C.c_int, [_P, _PP, _PP, _PP]),
    ("xt_cutoff", C.c_int, [_P, _P, _P, _P, _SZ, C.c_double, C.c_int, _PP, _PP, _PP, _DP, _DP]),
    ("xt_tensor_zeros_like", C.c_int, [_P, _PP]),
    ("xt_permute_adj", C.c_int, [_P, C.c_int, _IP, C.c_int, _PP]),
    ("xt_mul_bwd", C.c_int, [_P, _P, _P, _P, _P, C.c_int]),
    ("xt_svd_bwd", C.c_int, [_P, _P, _P, _P, _P, _P, _P]),
    ("xt_cutoff_bwd", C.c_int, [_P, _P, _P, _P, _P, _P]),
    ("xt_diag_bwd", C.c_int, [_P, _P, _P, C.c_int]),
    ("xt_trace_bwd", C.c_int, [_P, C.c_double]),
    ("xt_max_norm_bwd", C.c_int, [_P, _P, C.c_double]),
    ("xt_tqr", C.c_int, [_P, C.c_int, _PP, _PP]),
    ("xt_qr_bwd", C.c_int, [_P, _P, _P, _P, _P, _P]),
    ("xt_spec_create", C.c_int, [_PP]),
    ("xt_spec_destroy", C.c_int, [_P]),
    ("xt_spec_size", C.c_int, [_P, _IP]),
    ("xt_spec_entry", C.c_int, [_P, C.c_int, _IP, _IP, _DP, C.c_int]),
    ("xt_tsvd_spec", C.c_int, [_P, _SZ, C.c_double, _DP, _DP, _P, C.c_int, C.c_int, _PP, _PP, _PP]),
    ("xh_qr_bwd", C.c_int, [_P, _P, C.c_int64, _P, _P, _P, _P, _P, _P]),
    ("xt_teigh", C.c_int, [_P, _PP, _PP]),
    ("xt_adjoint", C.c_int, [_P, _PP]),
    ("xt_twist", C.c_int, [_P, C.c_int, _PP]),
    ("xt_scale", C.c_int, [_P, C.c_double, _PP]),
    ("xt_axpby", C.c_int, [_P, _P, C.c_double, C.c_double, _PP]),
    ("xt_diag_map", C.c_int, [_P, C.c_int, _PP]),
    ("xt_trace", C.c_int, [_P, _DP]),
    ("xt_squared_norm", C.c_int, [_P, _DP]),
    ("xt_max_norm", C.c_int, [_P, _DP]),
    ("xt_tape_begin", C.c_int, [_P]),
    ("xt_tape_pause", C.c_int, [_P, C.c_int]),
    ("xt_tape_seed", C.c_int, [_P, _P, _P, C.c_double]),
    ("xt_tape_seed_trace", C.c_int, [_P, _P, C.c_double]),
    ("xt_tape_backward", C.c_int, [_P]),
    ("xt_tape_adjoint", C.c_int, [_P, _P, _PP]),
    ("xt_tape_end", C.c_int, [_P]),
    ("xt_div_trace", C.c_int, [_P, _P, _PP]),
    ("xt_ctm_copy_env", C.c_int, [_P, _P]),
    ("xt_ctm_energy_grad", C.c_int, [_P, C.c_int, C.c_double, C.c_double, _DP, C.POINTER(C.c_void_p)]),
    ("xt_tensor_mexp", C.c_int, [_P, C.c_double, _PP]),
    ("xt_peps_create", C.c_int, [_P, C.POINTER(C.c_void_p), _SZ, _PP]),
    ("xt_peps_destroy", C.c_int, [_P]),
    ("xt_peps_tensor", C.c_int, [_P, C.c_int, C.c_int, _PP]),
    ("xt_tprop_create", C.c_int, [_P, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_double, _PP]),
    ("xt_tprop_destroy", C.c_int, [_P]),
    ("xt_tprop_step", C.c_int, [_P, _P]),
    ("xt_tprop_bond", C.c_int, [_P, _P, C.c_int, C.c_int, C.c_int]),
    ("xt_tprop_tensor", C.c_int, [_P, C.c_int, C.c_int, _PP]),
    ("xt_ctm_create", C.c_int, [_P, C.POINTER(C.c_void_p), _SZ, _PP]),
    ("xt_ctm_destroy", C.c_int, [_P]),
    ("xt_ctm_set_options", C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    ("xt_ctm_init", C.c_int, [_P]),
    ("xt_ctm_grow_all", C.c_int, [_P]),
    ("xt_ctm_compute_rdm", C.c_int, [_P]),
    ("xt_ctm_energy", C.c_int, [_P, C.c_double, C.c_double, _DP]),
    ("xt_ctm_tensor", C.c_int, [_P, C.c_int, C.c_int, _PP]),
    ("xt_ctm_predictor_stats", C.c_int, [_P, C.POINTER(C.c_long), C.POINTER(C.c_long)]),
    ("xt_ctm_svd_ms", C.c_int, [_P, _DP]),
    ("xt_ctm_set_rsvd", C.c_int, [_P, C.c_int, C.c_double, C.c_int, C.c_int]),
    ("xt_ctm_set_predictor", C.c_int, [_P, C.POINTER(C.c_float), C.c_int64, C.c_int, C.c_double]),
    ("xt_rank_features", C.c_int, [_DP, C.c_int, C.c_int, C.c_double, C.c_double, C.POINTER(C.c_float)]),
    ("xt_ctm_rsvd_stats", C.c_int, [_P, C.POINTER(C.c_long), C.POINTER(C.c_long), _DP]),
    ("xt_ctm_num_spectra", C.c_int, [_P, _IP]),
    ("xt_ctm_spectrum", C.c_int, [_P, C.c_int, _IP, _IP, _IP, _IP, _IP, _IP, _IP, _IP, _DP, C.c_int]),
    ("xt_ctm_clear_spectra", C.c_int, [_P]),
    ("xt_ctm_set_ranks", C.c_int, [_P, C.c_int, C.c_int, _IP, _IP, C.c_int]),
    ("xt_truncation_rule", C.c_int, [C.c_int, C.c_int, _IP, _DP, C.c_int64, _SZ, C.c_double, C.c_int, _IP, _IP, _IP, _DP, _DP]),
]

_lib = None
_finalizing = False


def _mark_finalizing():
    global _finalizing
    _finalizing = True


# Objects still alive when the interpreter exits are not torn down natively: their finalizers
# would run in an unspecified order, and destroying the device context (and its stream) while
# stream-ordered allocations made on it are still live leaves the HIP runtime's own exit
# handling with a dangling stream.  The process exit releases the device memory.
atexit.register(_mark_finalizing)


def native(name):
    """The library function `name` for a finalizer, or None once the interpreter is exiting."""
    if _finalizing or _lib is None:
        return None
    return getattr(_lib, name)


def lib():
    """Load the in-tree library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m xped_amd.build` (there is no CPU fallback)")
        L = C.CDLL(str(LIB_PATH), mode=C.RTLD_GLOBAL)
        for name, res, args in _SIGS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def declared_symbols():
    return [s[0] for s in _SIGS]


def check(rc, rt=None, what=""):
    if rc != XH_OK:
        L = lib()
        msg = L.xt_last_error(rt) if rt is not None else L.xt_last_error(None)
        msg = msg.decode() if msg else ""
        cls = XpedNoDevice if rc == XH_ENODEV else XpedError
        raise cls(rc, f"{what}: {msg}")


def device_count() -> int:
    n = C.c_int(0)
    lib().xh_device_count(C.byref(n))
    return n.value
