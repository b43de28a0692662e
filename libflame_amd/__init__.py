"""libflame_amd -- MI355X (gfx950) implementation of libflame's Cholesky / LU hot path.

Python host mirror of the C API in include/FLAME.h and include/FLA_hip.h,
through ctypes over the in-tree ``libflame_amd/libflame_amd.so``.  Function
names, argument meaning, return values and error behaviour follow libflame:
``FLA_Chol(uplo, A)`` returns FLA_SUCCESS (-1) or the 0-based index of the
first failing column, ``FLA_LU_piv(A, p)`` the 0-based index of the first
null pivot, and pivots are FLAME-native (0-based, relative to the row).

The library must be built (``python -c "import __graft_entry__ as g; g.build()"``
or ``make -C libflame_amd``); importing this package without it raises.
There is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libflame_amd.so")

# ---- constants (FLA_macro_defs.h values) ------------------------------------
FLA_SUCCESS, FLA_FAILURE = -1, -2
FLA_TL, FLA_TR, FLA_BL, FLA_BR = 11, 12, 21, 22
FLA_FLOAT, FLA_DOUBLE, FLA_COMPLEX, FLA_DOUBLE_COMPLEX, FLA_INT, FLA_CONSTANT = range(100, 106)
FLA_TOP, FLA_BOTTOM, FLA_LEFT, FLA_RIGHT = 200, 201, 210, 211
FLA_LOWER_TRIANGULAR, FLA_UPPER_TRIANGULAR = 300, 301
FLA_ZERO_MATRIX, FLA_FULL_MATRIX = 310, 311
FLA_NO_TRANSPOSE, FLA_TRANSPOSE, FLA_CONJ_TRANSPOSE = 400, 401, 402
FLA_UNIT_DIAG, FLA_NONUNIT_DIAG = 500, 501
FLA_NATIVE_PIVOTS, FLA_LAPACK_PIVOTS = 700, 701
FLA_NO_ERROR_CHECKING, FLA_MIN_ERROR_CHECKING, FLA_FULL_ERROR_CHECKING = 0, 1, 2
FLA_HIP_OK = 0


class FLA_Base_obj(C.Structure):
    _fields_ = [("datatype", C.c_int), ("elemtype", C.c_int), ("m", C.c_ulong), ("n", C.c_ulong),
                ("rs", C.c_ulong), ("cs", C.c_ulong), ("m_inner", C.c_ulong),
                ("n_inner", C.c_ulong), ("id", C.c_ulong), ("m_index", C.c_ulong),
                ("n_index", C.c_ulong), ("n_elem_alloc", C.c_ulong), ("buffer", C.c_void_p),
                ("buffer_info", C.c_int), ("uplo", C.c_int)]


class FLA_Obj(C.Structure):
    _fields_ = [("offm", C.c_ulong), ("offn", C.c_ulong), ("m", C.c_ulong), ("n", C.c_ulong),
                ("m_inner", C.c_ulong), ("n_inner", C.c_ulong),
                ("base", C.POINTER(FLA_Base_obj))]


_lib = None


def lib():
    """The loaded C library (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build() "
                               "(libflame_amd has no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        vp, ip, dp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)
        ulp = C.POINTER(C.c_ulong)
        O = FLA_Obj
        OP = C.POINTER(FLA_Obj)
        sig = {
            "FLA_Init": (None, []), "FLA_Finalize": (None, []),
            "FLA_Initialized": (C.c_int, []),
            "FLA_Check_error_level": (C.c_uint, []),
            "FLA_Check_error_level_set": (C.c_uint, [C.c_uint]),
            "FLA_Obj_create": (C.c_int, [C.c_int, C.c_ulong, C.c_ulong, C.c_ulong, C.c_ulong, OP]),
            "FLA_Obj_create_without_buffer": (C.c_int, [C.c_int, C.c_ulong, C.c_ulong, OP]),
            "FLA_Obj_attach_buffer": (C.c_int, [vp, C.c_ulong, C.c_ulong, OP]),
            "FLA_Obj_free": (C.c_int, [OP]), "FLA_Obj_free_without_buffer": (C.c_int, [OP]),
            "FLA_Obj_buffer_at_view": (vp, [O]),
            "FLA_Obj_is_device": (C.c_int, [O]),
            "FLA_Part_2x2": (C.c_int, [O, OP, OP, OP, OP, C.c_ulong, C.c_ulong, C.c_int]),
            "FLA_Chol": (C.c_int, [C.c_int, O]),
            "FLA_LU_piv": (C.c_int, [O, O]),
            "FLA_Apply_pivots": (C.c_int, [C.c_int, C.c_int, O, O]),
            "FLA_Chol_solve": (C.c_int, [C.c_int, O, O, O]),
            "FLA_LU_piv_solve": (C.c_int, [O, O, O, O]),
            "FLA_Shift_pivots_to": (C.c_int, [C.c_int, O]),
            "FLA_LU_nopiv": (C.c_int, [O]),
            "FLA_LU_nopiv_solve": (C.c_int, [O, O, O]),
            "FLA_Trinv": (C.c_int, [C.c_int, C.c_int, O]),
            "FLA_Ttmm": (C.c_int, [C.c_int, O]),
            "FLA_SPDinv": (C.c_int, [C.c_int, O]),
            "dtrtri_": (C.c_int, [C.c_char_p, C.c_char_p, ip, dp, ip, ip]),
            "dlauum_": (C.c_int, [C.c_char_p, ip, dp, ip, ip]),
            "dpotri_": (C.c_int, [C.c_char_p, ip, dp, ip, ip]),
            "dpotrf_": (C.c_int, [C.c_char_p, ip, dp, ip, ip]),
            "dgetrf_": (C.c_int, [ip, ip, dp, ip, ip, ip]),
            "dpotrs_": (C.c_int, [C.c_char_p, ip, ip, dp, ip, dp, ip, ip]),
            "dgetrs_": (C.c_int, [C.c_char_p, ip, ip, dp, ip, ip, dp, ip, ip]),
            "dlaswp_": (C.c_int, [ip, dp, ip, ip, ip, ip, ip]),
            "dposv_": (C.c_int, [C.c_char_p, ip, ip, dp, ip, dp, ip, ip]),
            "dgesv_": (C.c_int, [ip, ip, dp, ip, ip, dp, ip, ip]),
            "dsposv_": (C.c_int, [C.c_char_p, ip, ip, dp, ip, dp, ip, dp, ip, dp,
                                  C.POINTER(C.c_float), ip, ip]),
            "dsgesv_": (C.c_int, [ip, ip, dp, ip, ip, dp, ip, dp, ip, dp,
                                  C.POINTER(C.c_float), ip, ip]),
            "FLA_Obj_elemtype": (C.c_int, [O]),
            "FLA_Part_2x1": (C.c_int, [O, OP, OP, C.c_ulong, C.c_int]),
            "FLA_Part_1x2": (C.c_int, [O, OP, OP, C.c_ulong, C.c_int]),
            "FLASH_Obj_create": (C.c_int, [C.c_int, C.c_ulong, C.c_ulong, C.c_ulong, ulp, OP]),
            "FLASH_Obj_create_ext": (C.c_int, [C.c_int, C.c_ulong, C.c_ulong, C.c_ulong, ulp,
                                               ulp, OP]),
            "FLASH_Obj_create_without_buffer": (C.c_int, [C.c_int, C.c_ulong, C.c_ulong,
                                                          C.c_ulong, ulp, OP]),
            "FLASH_Obj_create_without_buffer_ext": (C.c_int, [C.c_int, C.c_ulong, C.c_ulong,
                                                              C.c_ulong, ulp, ulp, OP]),
            "FLASH_Obj_create_conf_to": (C.c_int, [C.c_int, O, OP]),
            "FLASH_Obj_create_hier_copy_of_flat": (C.c_int, [O, C.c_ulong, ulp, OP]),
            "FLASH_Obj_create_flat_copy_of_hier": (C.c_int, [O, OP]),
            "FLASH_Obj_free": (None, [OP]), "FLASH_Obj_free_without_buffer": (None, [OP]),
            "FLASH_Obj_extract_buffer": (vp, [O]),
            "FLASH_Obj_attach_buffer": (C.c_int, [vp, C.c_ulong, C.c_ulong, OP]),
            "FLASH_Obj_flatten": (C.c_int, [O, O]), "FLASH_Obj_hierarchify": (C.c_int, [O, O]),
            "FLASH_Copy_flat_to_hier": (C.c_int, [O, C.c_ulong, C.c_ulong, O]),
            "FLASH_Copy_hier_to_flat": (C.c_int, [C.c_ulong, C.c_ulong, O, O]),
            "FLASH_Copy": (C.c_int, [O, O]),
            "FLASH_Obj_depth": (C.c_ulong, [O]),
            "FLASH_Obj_blocksizes": (C.c_ulong, [O, ulp, ulp]),
            "FLASH_Obj_base_scalar_length": (C.c_ulong, [O]),
            "FLASH_Obj_base_scalar_width": (C.c_ulong, [O]),
            "FLASH_Obj_scalar_length": (C.c_ulong, [O]),
            "FLASH_Obj_scalar_width": (C.c_ulong, [O]),
            "FLASH_Obj_scalar_min_dim": (C.c_ulong, [O]),
            "FLASH_Obj_scalar_row_offset": (C.c_ulong, [O]),
            "FLASH_Obj_scalar_col_offset": (C.c_ulong, [O]),
            "FLASH_Obj_scalar_length_tl": (C.c_ulong, [O]),
            "FLASH_Obj_scalar_width_tl": (C.c_ulong, [O]),
            "FLASH_Chol": (C.c_int, [C.c_int, O]),
            "FLASH_Chol_solve": (C.c_int, [C.c_int, O, O, O]),
            "FLASH_LU_piv": (C.c_int, [O, O]),
            "FLASH_LU_piv_solve": (C.c_int, [O, O, O, O]),
            "FLASH_Apply_pivots": (C.c_int, [C.c_int, C.c_int, O, O]),
            "FLASH_LU_nopiv": (C.c_int, [O]),
            "FLASH_LU_nopiv_solve": (C.c_int, [O, O, O]),
            "FLASH_LU_find_zero_on_diagonal": (C.c_int, [O]),
            "FLASH_LU_incpiv": (C.c_int, [O, O, O]),
            "FLASH_LU_incpiv_create_hier_matrices": (C.c_int, [O, C.c_ulong, ulp, C.c_ulong, OP,
                                                               OP, OP]),
            "FLASH_LU_incpiv_determine_alg_blocksize": (C.c_ulong, [O]),
            "FLASH_FS_incpiv": (C.c_int, [O, O, O, O]),
            "FLASH_LU_incpiv_solve": (C.c_int, [O, O, O, O, O]),
            "FLASH_Trsm": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, O, O, O]),
            "fla_hip_dgetrf_incpiv": (C.c_int, [C.c_int, C.c_int, vp, C.c_int, C.c_int, C.c_int,
                                                vp, C.c_int, vp, C.c_int]),
            "fla_hip_dfs_incpiv": (C.c_int, [C.c_int, C.c_int, vp, C.c_int, C.c_int, C.c_int, vp,
                                             C.c_int, vp, C.c_int, C.c_int, vp, C.c_int]),
            "fla_hip_tile_copy": (C.c_int, [C.c_int, C.POINTER(C.c_longlong), vp, C.c_long,
                                            C.c_int, C.c_int]),
            "fla_hip_init": (C.c_int, []), "fla_hip_finalize": (C.c_int, []),
            "fla_hip_synchronize": (C.c_int, []),
            "fla_hip_set_stream": (C.c_int, [vp]),
            "fla_hip_get_stream": (vp, []),
            "fla_hip_malloc": (C.c_int, [C.POINTER(vp), C.c_size_t]),
            "fla_hip_free": (C.c_int, [vp]),
            "fla_hip_memcpy_h2d": (C.c_int, [vp, vp, C.c_size_t]),
            "fla_hip_memcpy_d2h": (C.c_int, [vp, vp, C.c_size_t]),
            "fla_hip_memcpy_d2d": (C.c_int, [vp, vp, C.c_size_t]),
            "fla_hip_memcpy2d": (C.c_int, [vp, C.c_size_t, vp, C.c_size_t, C.c_size_t,
                                           C.c_size_t, C.c_int]),
            "fla_hip_convert_d2s": (C.c_int, [C.c_int, C.c_int, vp, C.c_int, vp, C.c_int]),
            "fla_hip_batched_tune": (C.c_int, [C.c_int, C.c_int, ip, ip, ip, ip, dp]),
            "fla_hip_batched_set_config": (C.c_int, [C.c_int, C.c_int, C.c_int]),
            "fla_hip_batched_set_graph": (C.c_int, [C.c_int]),
            "fla_hip_set_gemm_config": (C.c_int, [C.c_int]),
            "fla_hip_set_gemm_group": (C.c_int, [C.c_int]),
            "fla_hip_set_reserved_cus": (C.c_int, [C.c_int]),
            "fla_hip_mp_set_panel": (C.c_int, [C.c_int]),
            "fla_hip_mp_set_reserved_cus": (C.c_int, [C.c_int]),
            "fla_hip_mp_set_blocksize": (C.c_int, [C.c_int]),
            "fla_hip_error_string": (C.c_char_p, [C.c_int]),
            "fla_hip_dgemm": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                        vp, C.c_int, vp, C.c_int, C.c_double, vp, C.c_int]),
            "fla_hip_dsyrk": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, vp,
                                        C.c_int, C.c_double, vp, C.c_int]),
            "fla_hip_dtrsm": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.c_double, vp, C.c_int, vp, C.c_int]),
            "fla_hip_dtrsm_noalloc": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                                C.c_int, C.c_double, vp, C.c_int, vp, C.c_int]),
            "fla_hip_dpotrf": (C.c_int, [C.c_int, C.c_int, vp, C.c_int, ip]),
            "fla_hip_dgetrf_nopiv": (C.c_int, [C.c_int, C.c_int, vp, C.c_int, ip]),
            "fla_hip_dtrtri": (C.c_int, [C.c_int, C.c_int, C.c_int, vp, C.c_int]),
            "fla_hip_dlauum": (C.c_int, [C.c_int, C.c_int, vp, C.c_int]),
            "fla_hip_dspdinv": (C.c_int, [C.c_int, C.c_int, vp, C.c_int, ip]),
            "fla_hip_dgetrf": (C.c_int, [C.c_int, C.c_int, vp, C.c_int, vp, ip]),
            "fla_hip_dlaswp": (C.c_int, [C.c_int, vp, C.c_int, C.c_int, C.c_int, vp, C.c_int]),
            "fla_hip_dlaswp_rows": (C.c_int, [C.c_int, C.c_int, vp, C.c_int, C.c_int, C.c_int, vp,
                                              C.c_int]),
            "fla_hip_shift_pivots": (C.c_int, [C.c_int, C.c_int, vp]),
            "fla_hip_dlange": (C.c_int, [C.c_int, C.c_int, C.c_int, vp, C.c_int, dp]),
            "fla_hip_dlansy": (C.c_int, [C.c_int, C.c_int, C.c_int, vp, C.c_int, dp]),
            "fla_hip_dlacpy": (C.c_int, [C.c_int, C.c_int, C.c_int, vp, C.c_int, vp, C.c_int]),
            "fla_hip_dlaset": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, vp,
                                         C.c_int]),
            "fla_hip_stats_enable": (C.c_int, [C.c_int]),
            "fla_hip_stats_reset": (C.c_int, []),
            "fla_hip_stats_get": (C.c_int, [C.c_int, dp, dp, C.POINTER(C.c_long)]),
            "fla_hip_set_blocksizes": (C.c_int, [C.c_int, C.c_int]),
            "fla_hip_get_blocksizes": (C.c_int, [ip, ip]),
            "fla_hip_chol_blocksize": (C.c_int, [C.c_int]),
            "fla_hip_gen_uniform": (C.c_int, [C.c_ulonglong, C.c_int, C.c_int, vp, C.c_int]),
            "fla_hip_gen_normal": (C.c_int, [C.c_ulonglong, C.c_int, C.c_int, vp, C.c_int]),
            "fla_hip_gram_lower": (C.c_int, [C.c_int, C.c_int, vp, C.c_int, C.c_double,
                                             C.c_double, vp, C.c_int]),
            "fla_hip_scale_rows_logspace": (C.c_int, [C.c_int, C.c_int, vp, C.c_int]),
            "fla_hip_dmatmul_check": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                                C.c_int, vp, C.c_int, vp, C.c_int, C.c_double,
                                                vp, C.c_int]),
            "fla_hip_gen_spd_reflectors": (C.c_int, [C.c_ulonglong, C.c_int, C.c_int,
                                                     C.c_double, vp, C.c_int]),
            "fla_hip_mp_chol": (C.c_int, [C.c_int, C.c_int, vp, C.c_int, C.c_int, vp, C.c_int,
                                          ip]),
            "fla_hip_dsposv": (C.c_int, [C.c_int, C.c_int, C.c_int, vp, C.c_int, vp, C.c_int,
                                         vp, C.c_int, C.POINTER(MpOpts), ip, ip]),
            "fla_hip_dsgesv": (C.c_int, [C.c_int, C.c_int, vp, C.c_int, vp, vp, C.c_int,
                                         vp, C.c_int, C.POINTER(MpOpts), ip, ip]),
            "fla_hip_mp_lu": (C.c_int, [C.c_int, vp, C.c_int, C.c_int, vp, C.c_int, vp, ip]),
            "fla_hip_schol_adjoint_batched": (C.c_int, [C.c_int, C.c_int, vp, vp, vp]),
            "fla_hip_sgp_rbf_lml_grad_batched": (C.c_int, [C.c_int, C.c_int, C.c_int, vp, vp, vp,
                                                           C.c_float, C.c_float, vp, vp, vp]),
            "fla_hip_gen_ge_reflectors": (C.c_int, [C.c_ulonglong, C.c_ulonglong, C.c_int,
                                                    C.c_int, C.c_double, vp, C.c_int]),
            "fla_hip_mp_last_stats": (C.c_int, [dp, C.c_int]),
            "fla_hip_dchol_adjoint": (C.c_int, [C.c_int, vp, C.c_int, vp, C.c_int, vp, C.c_int]),
            "fla_hip_dchol_solve_adjoint": (C.c_int, [C.c_int, C.c_int, vp, C.c_int, vp, C.c_int,
                                                      vp, C.c_int, vp, C.c_int, vp, C.c_int]),
            "fla_hip_gp_rbf_kernel": (C.c_int, [C.c_int, C.c_int, vp, C.c_int, C.c_double,
                                                C.c_double, C.c_double, vp, C.c_int]),
            "fla_hip_gp_rbf_grad": (C.c_int, [C.c_int, C.c_int, vp, C.c_int, C.c_double,
                                              C.c_double, vp, C.c_int, dp]),
            "fla_hip_gp_rbf_lml_grad": (C.c_int, [C.c_int, C.c_int, vp, C.c_int, vp, C.c_double,
                                                  C.c_double, C.c_double, dp, dp, ip]),
            "FLA_Chol_adjoint": (C.c_int, [C.c_int, O, O, O]),
            "FLA_Chol_solve_adjoint": (C.c_int, [C.c_int, O, O, O, O, O]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class MpOpts(C.Structure):
    """fla_hip_mp_opts (FLA_hip.h)."""
    _fields_ = [("precision", C.c_int), ("refinement", C.c_int), ("max_iter", C.c_int),
                ("gmres_restart", C.c_int), ("gmres_tol", C.c_double),
                ("stop_factor", C.c_double)]


FLA_MP_FP32, FLA_MP_BF16 = 0, 1
FLA_IR_CLASSIC, FLA_IR_GMRES = 0, 1
FLA_HIP_EARG = 9001
FLA_HIP_ERANGE = 9004
FLA_HIP_ENOINIT = 9005
FLA_HIP_ESCRATCH = 9006


def dsposv(uplo, A, B, X, precision=FLA_MP_FP32, refinement=FLA_IR_CLASSIC, max_iter=30,
           gmres_restart=20, gmres_tol=1e-4, stop_factor=0.0):
    """fla_hip_dsposv on device matrices; returns (iter, info, stats dict)."""
    o = MpOpts(precision, refinement, max_iter, gmres_restart, gmres_tol, stop_factor)
    it, info = C.c_int(), C.c_int()
    check(lib().fla_hip_dsposv(uplo, A.m, B.n, A.ptr, A.ld, B.ptr, B.ld, X.ptr, X.ld,
                               C.byref(o), C.byref(it), C.byref(info)), "fla_hip_dsposv")
    st = (C.c_double * 8)()
    check(lib().fla_hip_mp_last_stats(st, 8))
    keys = ["factor_ms", "refine_ms", "prec_solves", "matvecs", "anorm", "gmres_iters",
            "lp_info"]
    return it.value, info.value, {k: st[i] for i, k in enumerate(keys)}


def dsgesv(A, p, B, X, precision=FLA_MP_FP32, refinement=FLA_IR_CLASSIC, max_iter=30,
           gmres_restart=20, gmres_tol=1e-4, stop_factor=0.0):
    """fla_hip_dsgesv on device matrices (p: device int vector of n); returns
    (iter, info, stats dict)."""
    o = MpOpts(precision, refinement, max_iter, gmres_restart, gmres_tol, stop_factor)
    it, info = C.c_int(), C.c_int()
    check(lib().fla_hip_dsgesv(A.m, B.n, A.ptr, A.ld, p.ptr, B.ptr, B.ld, X.ptr, X.ld,
                               C.byref(o), C.byref(it), C.byref(info)), "fla_hip_dsgesv")
    st = (C.c_double * 8)()
    check(lib().fla_hip_mp_last_stats(st, 8))
    keys = ["factor_ms", "refine_ms", "prec_solves", "matvecs", "anorm", "gmres_iters",
            "lp_info"]
    return it.value, info.value, {k: st[i] for i, k in enumerate(keys)}


def check(status, what="libflame_amd call"):
    if status != FLA_HIP_OK:
        msg = lib().fla_hip_error_string(status).decode()
        raise RuntimeError(f"{what} failed: {msg} ({status})")


class DeviceBuffer:
    """Owned device allocation (fla_hip_malloc / fla_hip_free)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        check(lib().fla_hip_malloc(C.byref(self.ptr), max(self.nbytes, 8)), "fla_hip_malloc")

    def free(self):
        if self.ptr:
            lib().fla_hip_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def addr(self):
        return self.ptr.value


class DeviceMatrix:
    """Column-major m x n device matrix of float64 (or int32 vectors)."""

    def __init__(self, m, n, dtype=np.float64, ld=None):
        self.m, self.n = int(m), int(n)
        self.dtype = np.dtype(dtype)
        self.ld = int(ld) if ld is not None else max(1, self.m)
        self.buf = DeviceBuffer(self.ld * max(self.n, 1) * self.dtype.itemsize)

    @property
    def ptr(self):
        return self.buf.ptr

    @classmethod
    def from_host(cls, a, dtype=None):
        a = np.asarray(a)
        if a.ndim == 1:
            a = a.reshape(-1, 1)
        dtype = np.dtype(dtype or a.dtype)
        out = cls(a.shape[0], a.shape[1], dtype)
        out.upload(a)
        return out

    def upload(self, a):
        a = np.asfortranarray(np.asarray(a, dtype=self.dtype).reshape(self.m, self.n))
        if self.ld != self.m:
            raise ValueError("upload expects a packed matrix")
        check(lib().fla_hip_memcpy_h2d(self.ptr, a.ctypes.data, a.nbytes), "upload")

    def download(self):
        out = np.empty((self.ld, self.n), dtype=self.dtype, order="F")
        if out.nbytes:
            check(lib().fla_hip_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes), "download")
        return out[: self.m, :]

    def obj(self, datatype=None):
        """An FLA_Obj with this device buffer attached (the Python object keeps
        the base alive)."""
        dt = datatype or (FLA_INT if self.dtype == np.int32 else FLA_DOUBLE)
        return Obj.attach(self.ptr.value, self.m, self.n, dt, self.ld)


class Obj:
    """FLA_Obj created without buffer and attached to a given address."""

    def __init__(self):
        self.o = FLA_Obj()

    @classmethod
    def attach(cls, addr, m, n, datatype, ld):
        self = cls()
        lib().FLA_Obj_create_without_buffer(datatype, m, n, C.byref(self.o))
        lib().FLA_Obj_attach_buffer(C.c_void_p(addr), 1, max(1, ld), C.byref(self.o))
        return self

    def __del__(self):
        try:
            if self.o.base:
                lib().FLA_Obj_free_without_buffer(C.byref(self.o))
        except Exception:
            pass


class FlashObj:
    """A hierarchical (storage-by-blocks) object made by FLASH_Obj_create:
    m x n scalars in blocks of b (one size per level, or (b_m, b_n) lists),
    leaves in one device allocation."""

    def __init__(self, m, n, b, datatype=FLA_DOUBLE, b_n=None, buffer=True):
        bm = list(b) if isinstance(b, (list, tuple)) else [int(b)]
        bn = list(b_n) if b_n is not None else bm
        self.depth = len(bm)
        self.o = FLA_Obj()
        self.buffer = buffer
        am = (C.c_ulong * self.depth)(*bm)
        an = (C.c_ulong * self.depth)(*bn)
        fn = lib().FLASH_Obj_create_ext if buffer else lib().FLASH_Obj_create_without_buffer_ext
        r = fn(datatype, int(m), int(n), self.depth, am, an, C.byref(self.o))
        if r != FLA_SUCCESS:
            raise RuntimeError(f"FLASH_Obj_create failed ({r})")
        self.m, self.n, self.datatype = int(m), int(n), datatype

    @classmethod
    def wrap(cls, o, m, n, datatype=FLA_DOUBLE):
        """Takes ownership of a hierarchical FLA_Obj made by the library."""
        self = cls.__new__(cls)
        self.o, self.m, self.n, self.datatype, self.buffer = o, int(m), int(n), datatype, True
        self.depth = int(lib().FLASH_Obj_depth(o))
        return self

    @classmethod
    def from_host(cls, a, b, b_n=None):
        a = np.asarray(a)
        if a.ndim == 1:
            a = a.reshape(-1, 1)
        dt = FLA_INT if a.dtype == np.int32 else FLA_DOUBLE
        self = cls(a.shape[0], a.shape[1], b, dt, b_n)
        F = DeviceMatrix.from_host(a, np.int32 if dt == FLA_INT else np.float64)
        Fo = F.obj(dt)  # the wrapper must outlive the call (it frees the base)
        check_fla(lib().FLASH_Copy_flat_to_hier(Fo.o, 0, 0, self.o), "hierarchify")
        return self

    def download(self):
        dtype = np.int32 if self.datatype == FLA_INT else np.float64
        F = DeviceMatrix(self.m, self.n, dtype)
        Fo = F.obj(self.datatype)
        check_fla(lib().FLASH_Copy_hier_to_flat(0, 0, self.o, Fo.o), "flatten")
        return F.download()

    def free(self):
        if self.o.base:
            if self.buffer:
                lib().FLASH_Obj_free(C.byref(self.o))
            else:
                lib().FLASH_Obj_free_without_buffer(C.byref(self.o))

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def check_fla(r, what):
    if r != FLA_SUCCESS:
        raise RuntimeError(f"{what} failed ({r})")


def FLA_Init():
    lib().FLA_Init()


def FLA_Finalize():
    lib().FLA_Finalize()


def FLA_Chol(uplo, A):
    return lib().FLA_Chol(uplo, A.o if isinstance(A, Obj) else A)


def FLA_LU_nopiv(A):
    return lib().FLA_LU_nopiv(A.o if isinstance(A, Obj) else A)


def FLA_LU_nopiv_solve(A, B, X):
    return lib().FLA_LU_nopiv_solve(A.o, B.o, X.o)


def FLA_Trinv(uplo, diag, A):
    return lib().FLA_Trinv(uplo, diag, A.o if isinstance(A, Obj) else A)


def FLA_Ttmm(uplo, A):
    return lib().FLA_Ttmm(uplo, A.o if isinstance(A, Obj) else A)


def FLA_SPDinv(uplo, A):
    return lib().FLA_SPDinv(uplo, A.o if isinstance(A, Obj) else A)


def FLA_LU_piv(A, p):
    return lib().FLA_LU_piv(A.o if isinstance(A, Obj) else A, p.o if isinstance(p, Obj) else p)


def FLA_Apply_pivots(side, trans, p, A):
    return lib().FLA_Apply_pivots(side, trans, p.o, A.o)


def FLA_Chol_solve(uplo, A, B, X):
    return lib().FLA_Chol_solve(uplo, A.o, B.o, X.o)


def FLA_Chol_adjoint(uplo, L, Lbar, Abar):
    return lib().FLA_Chol_adjoint(uplo, L.o, Lbar.o, Abar.o)


def FLA_Chol_solve_adjoint(uplo, L, X, Xbar, Bbar, Abar):
    return lib().FLA_Chol_solve_adjoint(uplo, L.o, X.o, Xbar.o, Bbar.o, Abar.o)


def FLA_LU_piv_solve(A, p, B, X):
    return lib().FLA_LU_piv_solve(A.o, p.o, B.o, X.o)


def FLA_Shift_pivots_to(ptype, p):
    return lib().FLA_Shift_pivots_to(ptype, p.o)


def synchronize():
    check(lib().fla_hip_synchronize(), "synchronize")


def set_blocksizes(chol_nb=0, lu_nb=0):
    """0 keeps a block size, -1 restores its default."""
    check(lib().fla_hip_set_blocksizes(chol_nb, lu_nb), "set_blocksizes")


STATS_CHOL_UPDATE, STATS_LU_UPDATE, STATS_LU_PANEL, STATS_CHOL_PANEL = 0, 1, 2, 3
STATS_MP_UPDATE = 4
STATS_MP_LU_UPDATE = 5
STATS_LU_SWAP = 6


def stats_enable(on=True):
    lib().fla_hip_stats_enable(1 if on else 0)


def stats_reset():
    lib().fla_hip_stats_reset()


def stats_get(kind):
    """(algorithmic flops, milliseconds, launches) accumulated for `kind`."""
    f, ms, n = C.c_double(), C.c_double(), C.c_long()
    check(lib().fla_hip_stats_get(kind, C.byref(f), C.byref(ms), C.byref(n)), "stats_get")
    return f.value, ms.value, n.value


def get_blocksizes():
    a, b = C.c_int(), C.c_int()
    lib().fla_hip_get_blocksizes(C.byref(a), C.byref(b))
    return a.value, b.value
