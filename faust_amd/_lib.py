"""ctypes binding of libfaust_hip.so (the C ABI declared in include/faust_hip.h).

The library is built in-tree (faust_amd/libfaust_hip.so).  There is no CPU
fallback: importing this module without the built library raises.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfaust_hip.so")

FH_OK, FH_ERR_ARG, FH_ERR_HIP, FH_ERR_NUMERIC, FH_ERR_UNSUPPORTED = 0, 1, 2, 3, 4
FH_F64, FH_F32, FH_BF16 = 0, 1, 2
FH_MAT_DENSE, FH_MAT_CSR, FH_MAT_BSR = 0, 1, 2
# fh_faust_set_mul_mode values (faust_hip.h)
FH_MUL_AUTO, FH_MUL_PER_FACTOR, FH_MUL_FUSED = 0, 1, 2
FH_MUL_CSR_ROWS1, FH_MUL_CSR_ROWS2 = 32, 33
FH_BUTTERFLY_LEFT, FH_BUTTERFLY_RIGHT, FH_BUTTERFLY_BBTREE = 0, 1, 2


def FH_MUL_BSR_KERNEL(k):
    return 16 + k


PROX = {"sp": 0, "spcol": 1, "splin": 2, "splincol": 3, "supp": 4, "normcol": 5, "normlin": 6}


class FaustHipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[faust_hip error {code}] {msg}")
        self.code = code


class fh_constraint(C.Structure):
    _fields_ = [("name", C.c_int), ("k", C.c_int64), ("nrows", C.c_int), ("ncols", C.c_int),
                ("normalized", C.c_int), ("pos", C.c_int), ("supp", C.POINTER(C.c_double)), ("s", C.c_double)]


class fh_stop_crit(C.Structure):
    _fields_ = [("num_its", C.c_int), ("is_criterion_error", C.c_int), ("error_threshold", C.c_double),
                ("max_num_its", C.c_int)]


class fh_palm_params(C.Structure):
    _fields_ = [("is_update_way_R2L", C.c_int), ("factor_format", C.c_int), ("packing_RL", C.c_int),
                ("no_normalization", C.c_int), ("no_lambda", C.c_int), ("norm2_max_iter", C.c_int),
                ("norm2_threshold", C.c_double), ("constant_step_size", C.c_int), ("step_size", C.c_double),
                ("is_fact_side_left", C.c_int), ("record_traj", C.c_int), ("use_grad1", C.c_int),
                ("grad_calc_opt_mode", C.c_int)]


VP = C.c_void_p
_SIGS = {
    "fh_last_error": (C.c_char_p, []),
    "fh_version": (C.c_char_p, []),
    "fh_autotune_clear": (C.c_int, []),
    "fh_dev_count": (C.c_int, [C.POINTER(C.c_int)]),
    "fh_set_dev": (C.c_int, [C.c_int]),
    "fh_cur_dev": (C.c_int, [C.POINTER(C.c_int)]),
    "fh_stream_sync": (C.c_int, [VP]),
    "fh_mat_dense_create": (C.c_int, [C.c_int, C.c_int, C.c_int, VP, C.c_int, C.c_int, C.POINTER(VP)]),
    "fh_mat_csr_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, VP, VP, VP, C.c_int, C.POINTER(VP)]),
    "fh_mat_bsr_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, VP, VP, VP, C.c_int,
                                    C.POINTER(VP)]),
    "fh_mat_info": (C.c_int, [VP, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                              C.POINTER(C.c_int64)]),
    "fh_mat_tocpu": (C.c_int, [VP, VP, VP, VP]),
    "fh_mat_to_dense": (C.c_int, [VP, VP, VP]),
    "fh_mat_free": (C.c_int, [VP]),
    "fh_mat_mul_dense": (C.c_int, [VP, C.c_char, VP, C.c_int, C.c_int, VP, C.c_int, VP, VP, VP]),
    "fh_gemm": (C.c_int, [C.c_int, C.c_char, C.c_char, C.c_int, C.c_int, C.c_int, VP, VP, C.c_int, VP, C.c_int, VP,
                          VP, C.c_int, VP]),
    "fh_faust_multiply_noalloc": (C.c_int, [VP, C.c_char, VP, C.c_int, VP, VP]),
    "fh_faust_set_act_dtype": (C.c_int, [VP, C.c_int]),
    "fh_faust_get_act_dtype": (C.c_int, [VP, C.POINTER(C.c_int)]),
    "fh_faust_create": (C.c_int, [C.POINTER(VP)]),
    "fh_faust_push_back": (C.c_int, [VP, VP]),
    "fh_faust_size": (C.c_int, [VP, C.POINTER(C.c_int)]),
    "fh_faust_free": (C.c_int, [VP]),
    "fh_faust_factor": (C.c_int, [VP, C.c_int, C.POINTER(VP)]),
    "fh_butterfly": (C.c_int, [C.POINTER(C.c_double), C.c_int, C.c_int, C.POINTER(C.c_int32), C.c_int,
                               C.POINTER(VP)]),
    "fh_butterfly_device": (C.c_int, [VP, C.c_int, C.c_int64, C.c_int, C.c_int, C.POINTER(C.c_int32), C.c_int,
                                      C.POINTER(VP), VP]),
    "fh_faust_multiply": (C.c_int, [VP, C.c_char, VP, C.c_int, VP, VP]),
    "fh_faust_last_launches": (C.c_int, [VP, C.POINTER(C.c_int)]),
    "fh_faust_set_mul_mode": (C.c_int, [VP, C.c_int]),
    "fh_faust_get_mul_mode": (C.c_int, [VP, VP]),
    "fh_faust_get_product": (C.c_int, [VP, VP, VP]),
    "fh_faust_enable_timers": (C.c_int, [VP, C.c_int]),
    "fh_faust_timers": (C.c_int, [VP, VP, VP]),
    "fh_faust_spectral_norm": (C.c_int, [VP, C.c_int, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_int), VP]),
    "fh_prox": (C.c_int, [C.c_int, C.c_int, VP, C.c_int, C.c_int, C.c_int64, C.c_int, C.c_int, VP, VP]),
    "fh_prox_real": (C.c_int, [C.c_int, C.c_int, VP, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, VP]),
    "fh_norm_frob": (C.c_int, [C.c_int, VP, C.c_int64, C.POINTER(C.c_double), VP]),
    "fh_dot": (C.c_int, [C.c_int, VP, VP, C.c_int64, C.POINTER(C.c_double), VP]),
    "fh_palm4msa": (C.c_int, [VP, C.c_int, C.c_int, C.POINTER(fh_constraint), C.c_int, C.POINTER(fh_stop_crit),
                              C.POINTER(fh_palm_params), C.c_double, VP, C.POINTER(VP)]),
    "fh_hierarchical": (C.c_int, [VP, C.c_int, C.c_int, C.POINTER(fh_constraint), C.POINTER(fh_constraint), C.c_int,
                                  C.POINTER(fh_stop_crit), C.POINTER(fh_stop_crit), C.POINTER(fh_palm_params),
                                  C.POINTER(VP)]),
    "fh_palm4msa2016": (C.c_int, [VP, C.c_int, C.c_int, C.POINTER(fh_constraint), C.c_int, C.POINTER(fh_stop_crit),
                                  C.POINTER(fh_palm_params), C.c_double, VP, C.POINTER(VP)]),
    "fh_hierarchical2016": (C.c_int, [VP, C.c_int, C.c_int, C.POINTER(fh_constraint), C.POINTER(fh_constraint),
                                      C.c_int, C.POINTER(fh_stop_crit), C.POINTER(fh_stop_crit),
                                      C.POINTER(fh_palm_params), C.c_double, C.POINTER(VP)]),
    "fh_palm_result_nfacts": (C.c_int, [VP, C.POINTER(C.c_int)]),
    "fh_palm_result_lambda": (C.c_int, [VP, C.POINTER(C.c_double)]),
    "fh_palm_result_factor": (C.c_int, [VP, C.c_int, C.POINTER(VP)]),
    "fh_palm_result_traj": (C.c_int, [VP, VP, C.POINTER(C.c_int)]),
    "fh_palm_result_free": (C.c_int, [VP]),
    "fh_palm4msa_batched": (C.c_int, [C.c_int, C.c_int, VP, VP, VP, VP, VP, VP, VP, C.POINTER(VP)]),
    "fh_batch_result_lambda": (C.c_int, [VP, VP]),
    "fh_batch_result_relerr": (C.c_int, [VP, VP]),
    "fh_batch_result_factor": (C.c_int, [VP, C.c_int, C.c_int, VP]),
    "fh_batch_result_times": (C.c_int, [VP, VP]),
    "fh_batch_result_iters": (C.c_int, [VP, VP]),
    "fh_batch_result_free": (C.c_int, [VP]),
    "fh_mlp_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64, C.POINTER(VP)]),
    "fh_mlp_free": (C.c_int, [VP]),
    "fh_mlp_forward": (C.c_int, [VP, VP, C.c_int, VP, VP]),
    "fh_mlp_gradients": (C.c_int, [VP, VP, VP, C.c_int, C.POINTER(C.c_float), VP]),
    "fh_mlp_train_step": (C.c_int, [VP, VP, VP, C.c_int, C.c_float, C.POINTER(C.c_float), VP]),
    "fh_mlp_get": (C.c_int, [VP, C.c_int, C.c_int, VP, VP]),
    "fh_mlp_set": (C.c_int, [VP, C.c_int, VP, VP]),
}

_lib = None


def exported_names():
    return list(_SIGS.keys())


def lib():
    """Load libfaust_hip.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; "
                              f"g.build()' or make -C faust_amd/csrc)")
        l = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def check(rc):
    if rc != FH_OK:
        raise FaustHipError(rc, lib().fh_last_error().decode())
    return rc
