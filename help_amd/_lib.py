"""ctypes binding of include/help_mi355x.h (libhelp_mi355x.so, built in-tree).

Loading fails loudly when the library is missing: there is no Python or CPU
fallback for any compute entry point.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhelp_mi355x.so")

# status codes
HELP_OK = 0
HELP_ERR_ARG = 1
HELP_ERR_HIP = 2
HELP_ERR_NOT_PD = 3
HELP_ERR_STATE = 4
HELP_ERR_NOMEM = 5
HELP_ERR_NOCONV = 6
HELP_ERR_NOGPU = 7

MODE_DIRECT_FP64 = 0
MODE_DIRECT_FP32 = 1
MODE_PCG = 2
MODE_PCG_LEARNED = 3
FAULT_FP32_BREAKDOWN = 1
FAULT_REFINE_NOCONV = 2
FAULT_NET_NONFINITE = 4

GEN_RANDOM, GEN_MCF, GEN_STAIRCASE, GEN_ZIPF = 0, 1, 2, 3
ORDER_AMD, ORDER_ND = 0, 1

HSD_STATUS = {0: "OK", 1: "STALLED", 2: "ITER_LIMIT", 3: "UNKNOWN_ERROR"}


class Opts(C.Structure):
    _fields_ = [("mode", C.c_int32), ("beta", C.c_double), ("refine_max", C.c_int32),
                ("refine_tol", C.c_double), ("pcg_k", C.c_int32), ("pcg_max_iter", C.c_int32),
                ("pcg_tol", C.c_double), ("pcg_fp32_precond", C.c_int32),
                ("fault_flags", C.c_int32), ("verbose", C.c_int32), ("symbolic_only", C.c_int32),
                ("front_path", C.c_int32), ("profile", C.c_int32), ("refine_fp64", C.c_int32),
                ("inverse_solve", C.c_int32), ("ordering", C.c_int32), ("factor_graph", C.c_int32),
                ("refine_switch", C.c_double), ("ea_depth", C.c_int32)]


class SolveStats(C.Structure):
    _fields_ = [("refine_steps", C.c_int32), ("pcg_iters", C.c_int32), ("fallback", C.c_int32),
                ("rel_residual", C.c_double), ("ms_rhs", C.c_double), ("ms_solve", C.c_double),
                ("ms_back", C.c_double)]


class RenewStats(C.Structure):
    _fields_ = [("ms_assemble", C.c_double), ("ms_factor", C.c_double),
                ("ms_precond", C.c_double), ("n_tiny_pivots", C.c_int32), ("factor_bits", C.c_int32)]


class SymbolicInfo(C.Structure):
    _fields_ = [("nnz_adat_upper", C.c_int64), ("nnz_l", C.c_int64), ("nnz_l_exact", C.c_int64),
                ("n_supernodes", C.c_int64), ("n_levels", C.c_int64), ("max_front", C.c_int64),
                ("factor_flops", C.c_double), ("n_pairs", C.c_int64)]


class LP(C.Structure):
    _fields_ = [("m", C.c_int32), ("n", C.c_int32), ("nnz", C.c_int64), ("n_ub", C.c_int32),
                ("n_lb", C.c_int32), ("n_fr", C.c_int32), ("colptr", C.POINTER(C.c_int64)),
                ("rowind", C.POINTER(C.c_int32)), ("val", C.POINTER(C.c_double)),
                ("b", C.POINTER(C.c_double)), ("c", C.POINTER(C.c_double)),
                ("ub", C.POINTER(C.c_double)), ("c_intercept", C.c_double)]


class GenParams(C.Structure):
    _fields_ = [("kind", C.c_int32), ("seed", C.c_uint64), ("m", C.c_int32), ("nb", C.c_int32),
                ("k", C.c_int32), ("zipf_s", C.c_double), ("side", C.c_int32),
                ("commodities", C.c_int32), ("pairs", C.c_int32), ("periods", C.c_int32),
                ("block_rows", C.c_int32), ("block_cols", C.c_int32), ("block_k", C.c_int32),
                ("ub_cols", C.c_int32), ("free_cols", C.c_int32)]


CONSTRUCT_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(LP))
DESTRUCT_FN = C.CFUNCTYPE(C.c_int, C.c_void_p)
RENEW_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_double))
SOLVE_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                       C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double))


class Backend(C.Structure):
    _fields_ = [("ctx", C.c_void_p), ("construct", C.c_void_p), ("destruct", C.c_void_p),
                ("renew", C.c_void_p), ("solve", C.c_void_p)]


class HsdResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("iterations", C.c_int32), ("primal_obj", C.c_double),
                ("dual_obj", C.c_double), ("rel_gap", C.c_double), ("primal_infeas", C.c_double),
                ("dual_infeas", C.c_double), ("mu", C.c_double), ("seconds", C.c_double),
                ("seconds_renew", C.c_double), ("seconds_solve", C.c_double),
                ("seconds_construct", C.c_double), ("seconds_newton", C.c_double),
                ("seconds_direction", C.c_double), ("seconds_step", C.c_double),
                ("n_renew", C.c_int32), ("n_solve", C.c_int32), ("backend_error", C.c_int32),
                ("stall_reason", C.c_int32), ("newton_ratio", C.c_double)]


LOG_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int32, C.c_double, C.c_double, C.c_double,
                     C.c_double, C.c_double)

_P = C.c_void_p
_D = C.POINTER(C.c_double)
_I32 = C.POINTER(C.c_int32)
_I64 = C.POINTER(C.c_int64)

# every symbol the header declares, with its signature
SIGNATURES = {
    "help_opts_default": (None, [C.POINTER(Opts)]),
    "help_device_count": (C.c_int, []),
    "help_last_create_error": (C.c_char_p, []),
    "help_gpu_create": (C.c_int, [C.c_int32, C.c_int32, C.c_int64, _I64, _I32, _D,
                                  C.POINTER(Opts), C.POINTER(_P)]),
    "help_gpu_destroy": (C.c_int, [_P]),
    "help_gpu_renew": (C.c_int, [_P, _D, C.POINTER(RenewStats)]),
    "help_gpu_solve": (C.c_int, [_P, _D, _D, _D, _D, _D, C.POINTER(SolveStats)]),
    "help_gpu_solve2": (C.c_int, [_P, _D, _D, _D, _D, _D, _D, _D, _D, _D, C.POINTER(SolveStats),
                                  C.POINTER(SolveStats)]),
    "help_gpu_last_error": (C.c_char_p, [_P]),
    "help_gpu_symbolic_info": (C.c_int, [_P, C.POINTER(SymbolicInfo)]),
    "help_gpu_symbolic_arrays": (C.c_int, [_P, _I32, _I32, _I32]),
    "help_gpu_adat_pattern": (C.c_int, [_P, _I64, _I32]),
    "help_gpu_adat_values": (C.c_int, [_P, _D, _D]),
    "help_gpu_spmv": (C.c_int, [_P, C.c_int32, _D, _D]),
    "help_gpu_factor_column": (C.c_int, [_P, C.c_int32, _I32, _D, _I32]),
    "help_gpu_supernodes": (C.c_int, [_P, _I32, _I32, _I32, _I32]),
    "help_selftest_mfma": (C.c_int, [C.c_int32, _D, _D, _D]),
    "help_selftest_mfma_bf16": (C.c_int, [_D, _D, _D]),
    "help_generate": (C.c_int, [C.POINTER(GenParams), C.POINTER(LP)]),
    "help_lp_free": (None, [C.POINTER(LP)]),
    "help_write_mps": (C.c_int, [C.POINTER(LP), C.c_char_p, C.c_char_p]),
    "help_write_csc_bin": (C.c_int, [C.POINTER(LP), C.c_char_p]),
    "help_read_csc_bin": (C.c_int, [C.c_char_p, C.POINTER(LP)]),
    "help_hsd_solve": (C.c_int, [C.POINTER(LP), C.POINTER(Backend), C.c_int32, C.c_int32, LOG_FN,
                                 _P, C.POINTER(HsdResult)]),
    "help_hsd_solve_gpu": (C.c_int, [C.POINTER(LP), C.POINTER(Backend), C.c_int32, LOG_FN, _P,
                                     C.POINTER(HsdResult)]),
    "help_gpu_backend_new": (C.c_int, [C.POINTER(Opts), C.POINTER(Backend)]),
    "help_gpu_backend_free": (None, [C.POINTER(Backend)]),
    "help_gpu_backend_handle": (_P, [C.POINTER(Backend)]),
    "help_gpu_backend_kernel_stats": (C.c_int, [C.POINTER(Backend), C.c_int32, _D,
                                                C.POINTER(C.c_int64), _D]),
    "help_gpu_backend_kernel_stats_prec": (C.c_int, [C.POINTER(Backend), C.c_int32, C.c_int32, _D,
                                                     C.POINTER(C.c_int64), _D]),
    "help_gpu_backend_solve_stats": (C.c_int, [C.POINTER(Backend), C.POINTER(C.c_int64),
                                               C.POINTER(C.c_int64), _D]),
    "help_gpu_backend_solve2": (C.c_int, [C.POINTER(Backend), _D, _D, _D, _D, _D, _D, _D, _D, _D]),
    "help_mpnn_num_params": (C.c_int, []),
    "help_mpnn_new": (C.c_int, [C.c_int32, C.c_uint64, C.POINTER(C.c_void_p)]),
    "help_mpnn_free": (None, [_P]),
    "help_mpnn_get_params": (C.c_int, [_P, C.POINTER(C.c_float)]),
    "help_mpnn_set_params": (C.c_int, [_P, C.POINTER(C.c_float)]),
    "help_mpnn_predict": (C.c_int, [_P, _P, _D, _D, _D]),
    "help_mpnn_loss_grad": (C.c_int, [_P, _P, _D, _D, _D, _D, C.POINTER(C.c_float)]),
    "help_mpnn_adam_step": (C.c_int, [_P, C.c_double]),
    "help_gpu_set_warmstart": (C.c_int, [_P, _P]),
    "help_gpu_backend_set_warmstart": (C.c_int, [C.POINTER(Backend), _P]),
    "help_gpu_backend_pcg_iterations": (C.c_int64, [C.POINTER(Backend)]),
    "help_gpu_pcg_factor": (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32), _D, _D, _D]),
    "help_gpu_pcg_apply": (C.c_int, [_P, _D, _D]),
    "help_gpu_pcg_apply_noalloc": (C.c_int, [_P, C.c_void_p, C.c_void_p]),
    "help_gpu_profile": (C.c_int, [_P, C.c_int32]),
    "help_gpu_kernel_stats": (C.c_int, [_P, C.c_int32, _D, C.POINTER(C.c_int64), _D,
                                        C.POINTER(C.c_char_p)]),
    "help_device_synchronize": (C.c_int, []),
}

KERNEL_NAMES = ["k_small_fronts", "k_extend_add", "k_potrf_diag", "k_trsm_rows", "k_gemm", "k_fwd",
                "k_fwd_gemv", "k_bwd_gemv", "k_bwd", "k_adat", "k_rhs", "k_backsub", "pcg_build",
                "pcg_solve", "k_xinv", "k_fwd_xinv", "k_bwd_xinv"]
# work unit of each kernel class: flops for the factor kernels, HBM bytes otherwise
KERNEL_BOUND = ["mfma", "hbm", "mfma", "mfma", "mfma", "hbm", "hbm", "hbm", "hbm", "hbm", "hbm", "hbm",
                "hbm", "hbm", "mfma", "hbm", "hbm"]

_lib = None


def lib():
    """Load libhelp_mi355x.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def dptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_D)


def i32ptr(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_I32)


def i64ptr(a):
    assert a.dtype == np.int64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_I64)


def default_opts(**kw):
    o = Opts()
    lib().help_opts_default(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o
