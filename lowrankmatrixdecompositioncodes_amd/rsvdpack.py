"""Python view of librsvdpack's C ABI (include/rsvdpack.h, include/rsvdpack_ext.h).

The functions keep the reference names and argument meaning of RSVDPACK
(multi_core_mkl_code/rank_revealing_algorithms_intel_mkl.h): they take numpy
matrices, build the `mat` structs through the library's own matrix_new, call
the C entry point, and return numpy copies of the outputs (freed with the
library's matrix_delete).  Failures raise RsvdError with rsvd_last_error().

The library must be built (lowrankmatrixdecompositioncodes_amd/build.py);
importing this module without it raises -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

_LIB_PATH = Path(__file__).resolve().parent / "librsvdpack.so"


class RsvdError(RuntimeError):
    pass


class Mat(C.Structure):
    """MKL/mvf.h:18-21 -- {int nrows, ncols; double *d}, column-major."""
    _fields_ = [("nrows", C.c_int), ("ncols", C.c_int), ("d", C.POINTER(C.c_double))]


class Vec(C.Structure):
    """MKL/mvf.h:24-27 -- {int nrows; double *d}."""
    _fields_ = [("nrows", C.c_int), ("d", C.POINTER(C.c_double))]


MatP = C.POINTER(Mat)
VecP = C.POINTER(Vec)

# (name, restype, argtypes) of every exported entry point
_SIGS = {
    # L1 storage
    "matrix_new": (MatP, [C.c_int, C.c_int]),
    "vector_new": (VecP, [C.c_int]),
    "matrix_delete": (None, [MatP]),
    "vector_delete": (None, [VecP]),
    "matrix_set_element": (None, [MatP, C.c_int, C.c_int, C.c_double]),
    "matrix_get_element": (C.c_double, [MatP, C.c_int, C.c_int]),
    "vector_set_element": (None, [VecP, C.c_int, C.c_double]),
    "vector_get_element": (C.c_double, [VecP, C.c_int]),
    "matrix_load_from_binary_file": (MatP, [C.c_char_p]),
    "matrix_write_to_binary_file": (None, [MatP, C.c_char_p]),
    "matrix_print": (None, [MatP]),
    "vector_print": (None, [VecP]),
    "matrix_copy": (None, [MatP, MatP]),
    "matrix_build_transpose": (None, [MatP, MatP]),
    "matrix_sub": (None, [MatP, MatP]),
    "matrix_get_selected_columns": (None, [MatP, C.POINTER(C.c_int), MatP]),
    "matrix_set_selected_columns": (None, [MatP, C.POINTER(C.c_int), MatP]),
    "matrix_get_selected_rows": (None, [MatP, C.POINTER(C.c_int), MatP]),
    "matrix_set_selected_rows": (None, [MatP, C.POINTER(C.c_int), MatP]),
    "matrix_copy_symmetric": (None, [MatP, MatP]),
    "initialize_diagonal_matrix": (None, [MatP, VecP]),
    "invert_diagonal_matrix": (None, [MatP, MatP]),
    "resize_matrix_by_columns": (None, [C.POINTER(MatP), C.c_int]),
    "resize_matrix_by_columns_from_end": (None, [C.POINTER(MatP), C.c_int]),
    "resize_matrix_by_rows": (None, [C.POINTER(MatP), C.c_int]),
    "resize_matrix_by_rows_from_end": (None, [C.POINTER(MatP), C.c_int]),
    # L1 compute
    "initialize_random_matrix": (None, [MatP]),
    "matrix_matrix_mult": (None, [MatP, MatP, MatP]),
    "matrix_transpose_matrix_mult": (None, [MatP, MatP, MatP]),
    "matrix_matrix_transpose_mult": (None, [MatP, MatP, MatP]),
    "get_matrix_frobenius_norm": (C.c_double, [MatP]),
    "get_percent_error_between_two_mats": (C.c_double, [MatP, MatP]),
    "compute_evals_and_evecs_of_symm_matrix": (None, [MatP, VecP]),
    "compact_QR_factorization": (None, [MatP, MatP, MatP]),
    "QR_factorization_getQ": (None, [MatP, MatP]),
    "singular_value_decomposition": (None, [MatP, MatP, MatP, MatP]),
    "form_svd_product_matrix": (None, [MatP, MatP, MatP, MatP]),
    # L2
    "low_rank_svd_rand_decomp_fixed_rank": (
        None, [MatP, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int),
               C.POINTER(MatP), C.POINTER(MatP), C.POINTER(MatP)]),
    "low_rank_svd_blockrand_decomp_fixed_rank_or_prec": (
        None, [MatP, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int,
               C.POINTER(C.c_int), C.POINTER(MatP), C.POINTER(MatP), C.POINTER(MatP)]),
    "randQB_pb": (None, [MatP, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(MatP),
                         C.POINTER(MatP)]),
    "randQB_pb_new": (None, [MatP, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int,
                             C.POINTER(C.c_int), C.POINTER(MatP), C.POINTER(MatP)]),
    "use_low_rank_svd_for_approximation": (None, [MatP, MatP, MatP, MatP]),
    "use_QB_decomp_for_approximation": (None, [MatP, MatP, MatP]),
    "pivotedQR_mkl": (None, [MatP, C.POINTER(MatP), C.POINTER(MatP), C.POINTER(VecP)]),
    "id_blockrand_decomp_fixed_rank_or_prec": (
        None, [MatP, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int,
               C.POINTER(C.c_int), C.POINTER(VecP), C.POINTER(MatP)]),
    "id_two_sided_blockrand_decomp_fixed_rank_or_prec": (
        None, [MatP, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int,
               C.POINTER(C.c_int), C.POINTER(VecP), C.POINTER(VecP), C.POINTER(MatP),
               C.POINTER(MatP)]),
    "cur_blockrand_decomp_fixed_rank_or_prec": (
        None, [MatP, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int,
               C.POINTER(C.c_int), C.POINTER(MatP), C.POINTER(MatP), C.POINTER(MatP)]),
    "form_cur_product_matrix": (None, [MatP, MatP, MatP, MatP]),
    "use_id_decomp_for_approximation": (None, [MatP, MatP, VecP, C.c_int]),
    "use_id_two_sided_decomp_for_approximation": (None, [MatP, MatP, MatP, VecP, VecP, C.c_int]),
    "use_cur_decomp_for_approximation": (None, [MatP, MatP, MatP, MatP]),
    # extensions
    "rsvd_last_error": (C.c_char_p, []),
    "rsvd_init": (C.c_int, []),
    "rsvd_version": (C.c_char_p, []),
    "rsvd_set_seed": (None, [C.c_ulonglong]),
    "rsvd_use_time_seed": (None, []),
    "rsvd_set_omega": (C.c_int, [C.POINTER(C.c_double), C.c_longlong]),
    "rsvd_gaussian": (C.c_int, [C.c_ulonglong, C.c_longlong, C.c_longlong,
                                C.POINTER(C.c_double)]),
    "rsvd_dev_alloc": (C.c_void_p, [C.c_size_t]),
    "rsvd_dev_free": (None, [C.c_void_p]),
    "rsvd_memcpy_h2d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "rsvd_memcpy_d2h": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "rsvd_memset_dev": (C.c_int, [C.c_void_p, C.c_int, C.c_size_t]),
    "rsvd_dev_gaussian": (C.c_int, [C.c_ulonglong, C.c_longlong, C.c_longlong, C.c_void_p,
                                    C.c_longlong]),
    "rsvd_sync": (C.c_int, []),
    "rsvd_release_cache": (None, []),
    "rsvd_dev_fixed_rank": (
        C.c_int, [C.c_void_p, C.c_longlong, C.c_longlong, C.c_longlong, C.c_int, C.c_int,
                  C.c_int, C.c_int, C.c_int, C.c_ulonglong, C.c_void_p, C.c_void_p,
                  C.c_longlong, C.c_void_p, C.c_void_p, C.c_longlong]),
    "rsvd_dev_fixed_rank_noalloc": (
        C.c_int, [C.c_void_p, C.c_longlong, C.c_longlong, C.c_longlong, C.c_int, C.c_int,
                  C.c_int, C.c_int, C.c_int, C.c_ulonglong, C.c_void_p, C.c_void_p,
                  C.c_longlong, C.c_void_p, C.c_void_p, C.c_longlong]),
    "rsvd_binary_file_shape": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_longlong),
                                         C.POINTER(C.c_longlong)]),
    "rsvd_dev_load_binary": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_void_p, C.c_longlong]),
    "rsvd_dev_write_binary": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_void_p, C.c_longlong,
                                        C.c_longlong, C.c_longlong]),
    "rsvd_qb_max_cols": (C.c_longlong, [C.c_longlong, C.c_longlong, C.c_int, C.c_int, C.c_int]),
    "rsvd_dev_randqb": (
        C.c_int, [C.c_void_p, C.c_longlong, C.c_longlong, C.c_longlong, C.c_int, C.c_int,
                  C.c_double, C.c_int, C.c_int, C.c_int, C.c_ulonglong, C.c_void_p, C.c_void_p,
                  C.c_longlong, C.c_void_p, C.c_longlong, C.POINTER(C.c_int),
                  C.POINTER(C.c_double)]),
    "rsvd_dev_blockrand": (
        C.c_int, [C.c_void_p, C.c_longlong, C.c_longlong, C.c_longlong, C.c_int, C.c_int,
                  C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.c_ulonglong, C.c_void_p,
                  C.POINTER(C.c_int), C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                  C.POINTER(C.c_void_p)]),
    "rsvd_dev_copy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_longlong]),
    "rsvd_dev_fixed_rank_batched": (C.c_int, [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                              C.c_int, C.c_ulonglong, C.c_void_p, C.c_void_p,
                                              C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                              C.c_void_p, C.c_void_p, C.POINTER(C.c_int)]),
    "rsvd_batched_last_config": (C.c_char_p, []),
    "rsvd_mlp_num_params": (C.c_int, []),
    "rsvd_dev_mlp_init": (C.c_int, [C.c_ulonglong, C.c_void_p]),
    "rsvd_dev_mlp_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    "rsvd_dev_mlp_grad": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_float,
                                    C.c_int, C.c_void_p, C.c_void_p]),
    "rsvd_dev_mlp_train": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                     C.c_float, C.c_int, C.c_int, C.c_float, C.c_ulonglong,
                                     C.c_void_p]),
    "rsvd_dev_sketch_features": (C.c_int, [C.c_void_p, C.c_void_p, C.c_longlong, C.c_longlong,
                                           C.c_longlong, C.c_ulonglong, C.c_void_p]),
    "rsvd_dev_learned_lowrank": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_longlong,
                                           C.c_longlong, C.c_longlong, C.c_double, C.c_int,
                                           C.c_double, C.c_int, C.c_int, C.c_ulonglong,
                                           C.POINTER(C.c_int), C.POINTER(C.c_void_p),
                                           C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                           C.POINTER(C.c_int), C.POINTER(C.c_int),
                                           C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "rsvd_dev_batch_test_matrices": (C.c_int, [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                               C.c_ulonglong, C.c_void_p, C.c_void_p]),
    "rsvd_dev_gemm": (C.c_int, [C.c_int, C.c_char, C.c_char, C.c_longlong, C.c_longlong,
                                C.c_longlong, C.c_double, C.c_void_p, C.c_longlong, C.c_void_p,
                                C.c_longlong, C.c_double, C.c_void_p, C.c_longlong]),
    "rsvd_dev_sym_eig_psd": (C.c_int, [C.c_longlong, C.c_void_p, C.c_longlong, C.c_void_p,
                                       C.c_void_p, C.c_longlong, C.POINTER(C.c_int)]),
    "rsvd_dev_householder_qr": (C.c_int, [C.c_void_p, C.c_longlong, C.c_longlong, C.c_longlong,
                                          C.c_void_p]),
    "rsvd_dev_lowrank_test_matrix": (C.c_int, [C.c_longlong, C.c_longlong, C.c_int, C.c_void_p,
                                               C.c_double, C.c_ulonglong, C.c_void_p,
                                               C.c_longlong, C.c_void_p, C.c_longlong]),
    "rsvd_dev_fixed_rank_mp": (C.c_int, [C.c_void_p, C.c_longlong, C.c_longlong, C.c_longlong,
                                         C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                         C.c_ulonglong, C.c_void_p, C.c_int, C.c_void_p,
                                         C.c_longlong, C.c_void_p, C.c_void_p, C.c_longlong]),
    "rsvd_dev_randqb_mp": (
        C.c_int, [C.c_void_p, C.c_longlong, C.c_longlong, C.c_longlong, C.c_int, C.c_int,
                  C.c_double, C.c_int, C.c_int, C.c_int, C.c_ulonglong, C.c_void_p, C.c_int,
                  C.c_void_p, C.c_longlong, C.c_void_p, C.c_longlong, C.POINTER(C.c_int),
                  C.POINTER(C.c_double)]),
    "rsvd_dev_blockrand_mp": (
        C.c_int, [C.c_void_p, C.c_longlong, C.c_longlong, C.c_longlong, C.c_int, C.c_int,
                  C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.c_ulonglong, C.c_void_p,
                  C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                  C.POINTER(C.c_void_p)]),
    "rsvd_dev_lowrank_residual": (C.c_int, [C.c_void_p, C.c_longlong, C.c_longlong,
                                            C.c_longlong, C.c_longlong, C.c_void_p,
                                            C.c_longlong, C.c_void_p, C.c_void_p, C.c_longlong,
                                            C.c_int, C.POINTER(C.c_double)]),
    "rsvd_timer_start": (C.c_int, []),
    "rsvd_timer_stop_ms": (C.c_double, []),
    "rsvd_kstats_enable": (None, [C.c_int]),
    "rsvd_set_fp64_vendor_gemm": (None, [C.c_int]),
    "rsvd_kstats_reset": (None, []),
    "rsvd_kstats_select": (None, [C.c_char_p]),
    "rsvd_kstats_get": (C.c_int, [C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_longlong)]),
}

_lib = None


def lib_path() -> Path:
    return _LIB_PATH


def load() -> C.CDLL:
    """Load librsvdpack.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            raise RsvdError(f"{_LIB_PATH} is missing: run lowrankmatrixdecompositioncodes_amd/build.py")
        lib = C.CDLL(str(_LIB_PATH))
        for name, (res, args) in _SIGS.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def last_error() -> str:
    return (load().rsvd_last_error() or b"").decode()


def _check(rc: int = 0, what: str = "") -> None:
    msg = last_error()
    if rc or msg:
        raise RsvdError(f"{what}: {msg or 'rc=%d' % rc}")


# ------------------------------------------------------------------ mat glue
def to_mat(a: np.ndarray) -> MatP:
    """New library-owned `mat` holding a copy of the 2-D array a."""
    lib = load()
    a = np.asarray(a, dtype=np.float64)
    if a.ndim != 2:
        raise ValueError("expected a 2-D array")
    m, n = a.shape
    M = lib.matrix_new(m, n)
    if m * n:
        buf = np.asfortranarray(a)
        C.memmove(M.contents.d, buf.ctypes.data, 8 * m * n)
    return M


def from_mat(M: MatP, free: bool = True) -> np.ndarray:
    lib = load()
    if not M:
        raise RsvdError("NULL matrix: " + last_error())
    m, n = M.contents.nrows, M.contents.ncols
    out = np.empty((m, n), dtype=np.float64, order="F")
    if m * n:
        C.memmove(out.ctypes.data, M.contents.d, 8 * m * n)
    if free:
        lib.matrix_delete(M)
    return np.ascontiguousarray(out)


def empty_mat(m: int, n: int) -> MatP:
    return load().matrix_new(m, n)


# ------------------------------------------------------------------ configuration
def set_seed(seed: int) -> None:
    load().rsvd_set_seed(C.c_ulonglong(seed))


def use_time_seed() -> None:
    load().rsvd_use_time_seed()


def set_omega(omega: np.ndarray | None) -> None:
    lib = load()
    if omega is None:
        lib.rsvd_set_omega(None, 0)
        return
    buf = np.asfortranarray(np.asarray(omega, dtype=np.float64))
    _check(lib.rsvd_set_omega(buf.ctypes.data_as(C.POINTER(C.c_double)), buf.size),
           "rsvd_set_omega")


def gaussian(seed: int, offset: int, count: int) -> np.ndarray:
    out = np.empty(count, dtype=np.float64)
    lib = load()
    _check(lib.rsvd_gaussian(seed, offset, count, out.ctypes.data_as(C.POINTER(C.c_double))),
           "rsvd_gaussian")
    return out


# ------------------------------------------------------------------ L1 compute
def _gemm(fn: str, A: np.ndarray, B: np.ndarray, shape: tuple[int, int]) -> np.ndarray:
    lib = load()
    Am, Bm, Cm = to_mat(A), to_mat(B), empty_mat(*shape)
    try:
        getattr(lib, fn)(Am, Bm, Cm)
        _check(0, fn)
        return from_mat(Cm, free=False)
    finally:
        lib.matrix_delete(Am)
        lib.matrix_delete(Bm)
        lib.matrix_delete(Cm)


def matrix_matrix_mult(A, B):
    return _gemm("matrix_matrix_mult", A, B, (A.shape[0], B.shape[1]))


def matrix_transpose_matrix_mult(A, B):
    return _gemm("matrix_transpose_matrix_mult", A, B, (A.shape[1], B.shape[1]))


def matrix_matrix_transpose_mult(A, B):
    return _gemm("matrix_matrix_transpose_mult", A, B, (A.shape[0], B.shape[0]))


def QR_factorization_getQ(M):
    lib = load()
    Mm, Qm = to_mat(M), empty_mat(*M.shape)
    try:
        lib.QR_factorization_getQ(Mm, Qm)
        _check(0, "QR_factorization_getQ")
        return from_mat(Qm, free=False)
    finally:
        lib.matrix_delete(Mm)
        lib.matrix_delete(Qm)


def compact_QR_factorization(M):
    lib = load()
    m, n = M.shape
    k = min(m, n)
    Mm, Qm, Rm = to_mat(M), empty_mat(m, n), empty_mat(k, k)
    try:
        lib.compact_QR_factorization(Mm, Qm, Rm)
        _check(0, "compact_QR_factorization")
        return from_mat(Qm, free=False), from_mat(Rm, free=False)
    finally:
        for x in (Mm, Qm, Rm):
            lib.matrix_delete(x)


def compute_evals_and_evecs_of_symm_matrix(S):
    lib = load()
    n = S.shape[0]
    Sm = to_mat(S)
    ev = lib.vector_new(n)
    try:
        lib.compute_evals_and_evecs_of_symm_matrix(Sm, ev)
        _check(0, "compute_evals_and_evecs_of_symm_matrix")
        vals = np.array([ev.contents.d[i] for i in range(n)])
        return vals, from_mat(Sm, free=False)
    finally:
        lib.matrix_delete(Sm)
        lib.vector_delete(ev)


def singular_value_decomposition(M):
    lib = load()
    m, n = M.shape
    k = min(m, n)
    Mm, Um, Sm, Vtm = to_mat(M), empty_mat(m, k), empty_mat(k, k), empty_mat(k, n)
    try:
        lib.singular_value_decomposition(Mm, Um, Sm, Vtm)
        _check(0, "singular_value_decomposition")
        return from_mat(Um, False), from_mat(Sm, False), from_mat(Vtm, False)
    finally:
        for x in (Mm, Um, Sm, Vtm):
            lib.matrix_delete(x)


def get_matrix_frobenius_norm(M) -> float:
    lib = load()
    Mm = to_mat(M)
    try:
        v = lib.get_matrix_frobenius_norm(Mm)
        _check(0, "get_matrix_frobenius_norm")
        return v
    finally:
        lib.matrix_delete(Mm)


def get_percent_error_between_two_mats(A, B) -> float:
    lib = load()
    Am, Bm = to_mat(A), to_mat(B)
    try:
        v = lib.get_percent_error_between_two_mats(Am, Bm)
        _check(0, "get_percent_error_between_two_mats")
        return v
    finally:
        lib.matrix_delete(Am)
        lib.matrix_delete(Bm)


def initialize_random_matrix(m: int, n: int) -> np.ndarray:
    lib = load()
    Mm = empty_mat(m, n)
    try:
        lib.initialize_random_matrix(Mm)
        _check(0, "initialize_random_matrix")
        return from_mat(Mm, free=False)
    finally:
        lib.matrix_delete(Mm)


# ------------------------------------------------------------------ L2
def low_rank_svd_rand_decomp_fixed_rank(M, k, p, vnum, q, s):
    """MKL/rra.c:73-234.  Returns (U, S, V) as numpy arrays (S is k x k)."""
    lib = load()
    Mm = to_mat(M)
    U, S, V = MatP(), MatP(), MatP()
    frank = C.c_int(-12345)
    try:
        lib.low_rank_svd_rand_decomp_fixed_rank(Mm, k, p, vnum, q, s, C.byref(frank),
                                                C.byref(U), C.byref(S), C.byref(V))
        _check(0, "low_rank_svd_rand_decomp_fixed_rank")
        return from_mat(U), from_mat(S), from_mat(V)
    finally:
        lib.matrix_delete(Mm)


def low_rank_svd_blockrand_decomp_fixed_rank_or_prec(M, k, p, TOL, vnum, kstep, q, s):
    """MKL/rra.c:239-381.  Returns (U, S, V, frank)."""
    lib = load()
    Mm = to_mat(M)
    U, S, V = MatP(), MatP(), MatP()
    frank = C.c_int(0)
    try:
        lib.low_rank_svd_blockrand_decomp_fixed_rank_or_prec(
            Mm, k, p, TOL, vnum, kstep, q, s, C.byref(frank), C.byref(U), C.byref(S),
            C.byref(V))
        _check(0, "low_rank_svd_blockrand_decomp_fixed_rank_or_prec")
        return from_mat(U), from_mat(S), from_mat(V), frank.value
    finally:
        lib.matrix_delete(Mm)


def randQB_pb(M, kstep, nstep, q, s):
    """MKL/rra.c:1425-1569.  Returns (Q, B)."""
    lib = load()
    Mm = to_mat(M)
    Q, B = MatP(), MatP()
    try:
        lib.randQB_pb(Mm, kstep, nstep, q, s, C.byref(Q), C.byref(B))
        _check(0, "randQB_pb")
        return from_mat(Q), from_mat(B)
    finally:
        lib.matrix_delete(Mm)


def randQB_pb_new(M, kstep, nstep, TOL, q, s):
    """MKL/rra.c:1576-1801.  Returns (Q, B, frank)."""
    lib = load()
    Mm = to_mat(M)
    Q, B = MatP(), MatP()
    frank = C.c_int(0)
    try:
        lib.randQB_pb_new(Mm, kstep, nstep, TOL, q, s, C.byref(frank), C.byref(Q), C.byref(B))
        _check(0, "randQB_pb_new")
        return from_mat(Q), from_mat(B), frank.value
    finally:
        lib.matrix_delete(Mm)


# ------------------------------------------------------------------ device arrays
def _from_vec(v) -> np.ndarray:
    n = v.contents.nrows
    out = np.array([v.contents.d[i] for i in range(n)], dtype=np.float64)
    load().vector_delete(v)
    return out


def pivotedQR_mkl(M):
    """Column-pivoted QR; returns (Q, R, I) with M[:, I] = Q R."""
    lib = load()
    Mm = to_mat(M)
    Q, R, I = MatP(), MatP(), VecP()
    lib.pivotedQR_mkl(Mm, C.byref(Q), C.byref(R), C.byref(I))
    lib.matrix_delete(Mm)
    if not Q:
        raise RsvdError("pivotedQR_mkl: " + last_error())
    return from_mat(Q), from_mat(R), _from_vec(I).astype(np.int64)


def id_blockrand_decomp_fixed_rank_or_prec(M, k, p, TOL, kstep, q, s):
    """Column ID; returns (I as int64 0-based indices, T, frank)."""
    lib = load()
    Mm = to_mat(M)
    fr = C.c_int(0)
    I, T = VecP(), MatP()
    lib.id_blockrand_decomp_fixed_rank_or_prec(Mm, k, p, TOL, kstep, q, s, C.byref(fr),
                                               C.byref(I), C.byref(T))
    lib.matrix_delete(Mm)
    if not I or not T:
        raise RsvdError("id_blockrand_decomp_fixed_rank_or_prec: " + last_error())
    return _from_vec(I).astype(np.int64), from_mat(T), fr.value


def id_two_sided_blockrand_decomp_fixed_rank_or_prec(M, k, p, TOL, kstep, q, s):
    """Two-sided ID; returns (Icol, Irow, T, S, frank)."""
    lib = load()
    Mm = to_mat(M)
    fr = C.c_int(0)
    Ic, Ir, T, S = VecP(), VecP(), MatP(), MatP()
    lib.id_two_sided_blockrand_decomp_fixed_rank_or_prec(Mm, k, p, TOL, kstep, q, s, C.byref(fr),
                                                         C.byref(Ic), C.byref(Ir), C.byref(T),
                                                         C.byref(S))
    lib.matrix_delete(Mm)
    if not Ic or not T:
        raise RsvdError("id_two_sided_blockrand_decomp_fixed_rank_or_prec: " + last_error())
    return (_from_vec(Ic).astype(np.int64), _from_vec(Ir).astype(np.int64), from_mat(T),
            from_mat(S), fr.value)


def cur_blockrand_decomp_fixed_rank_or_prec(M, k, p, TOL, kstep, q, s):
    """CUR; returns (C, U, R, frank)."""
    lib = load()
    Mm = to_mat(M)
    fr = C.c_int(0)
    Cm, Um, Rm = MatP(), MatP(), MatP()
    lib.cur_blockrand_decomp_fixed_rank_or_prec(Mm, k, p, TOL, kstep, q, s, C.byref(fr),
                                                C.byref(Cm), C.byref(Um), C.byref(Rm))
    lib.matrix_delete(Mm)
    if not Cm:
        raise RsvdError("cur_blockrand_decomp_fixed_rank_or_prec: " + last_error())
    return from_mat(Cm), from_mat(Um), from_mat(Rm), fr.value


class DeviceMatrix:
    """A column-major matrix in HBM, allocated through rsvd_dev_alloc."""

    def __init__(self, rows: int, cols: int, dtype=np.float64):
        self.rows, self.cols = int(rows), int(cols)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.rows * self.cols * self.dtype.itemsize
        p = load().rsvd_dev_alloc(max(self.nbytes, 1))
        if not p:
            raise RsvdError("rsvd_dev_alloc: " + last_error())
        self.ptr = p

    @classmethod
    def from_host(cls, a: np.ndarray, dtype=None) -> "DeviceMatrix":
        a = np.asarray(a, dtype=dtype or a.dtype)
        d = cls(a.shape[0], a.shape[1], a.dtype)
        buf = np.asfortranarray(a)
        if d.nbytes:
            _check(load().rsvd_memcpy_h2d(d.ptr, buf.ctypes.data, d.nbytes), "h2d")
        return d

    def to_host(self) -> np.ndarray:
        out = np.empty((self.rows, self.cols), dtype=self.dtype, order="F")
        if self.nbytes:
            _check(load().rsvd_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes), "d2h")
        return out

    def free(self) -> None:
        if self.ptr:
            load().rsvd_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def dev_fixed_rank(A: DeviceMatrix, k, p, vnum, q, s, seed, omega: DeviceMatrix | None = None):
    """Device-resident fixed-rank rSVD (FP64); returns host (U, svals, V)."""
    lib = load()
    m, n = A.rows, A.cols
    U, Sv, V = DeviceMatrix(m, k), DeviceMatrix(k, 1), DeviceMatrix(n, k)
    rc = lib.rsvd_dev_fixed_rank(A.ptr, m, n, m, k, p, vnum, q, s, seed,
                                 omega.ptr if omega is not None else None,
                                 U.ptr, m, Sv.ptr, V.ptr, n)
    _check(rc, "rsvd_dev_fixed_rank")
    return U.to_host(), Sv.to_host()[:, 0], V.to_host()


def dev_gemm(ta: str, tb: str, A: DeviceMatrix, B: DeviceMatrix, C_: DeviceMatrix,
             alpha: float = 1.0, beta: float = 0.0, mode: str | None = None) -> None:
    """C = alpha*op(A)*op(B) + beta*C on device matrices of one dtype (f64 / f32).

    mode (float32 only): None -> fp32 matrix cores, "bf16" -> bf16 products with
    fp32 accumulation, "bf16x3" -> split-bf16 products.
    """
    M = C_.rows
    N = C_.cols
    K = A.rows if ta == "T" else A.cols
    dt = {np.dtype(np.float64): 0, np.dtype(np.float32): 1}[A.dtype]
    if mode is not None:
        dt = {"bf16": 2, "bf16x3": 3}[mode]
    rc = load().rsvd_dev_gemm(dt, ta.encode(), tb.encode(), M, N, K, alpha, A.ptr, A.rows,
                              B.ptr, B.rows, beta, C_.ptr, C_.rows)
    _check(rc, "rsvd_dev_gemm")


def kstats(name: str) -> tuple[float, int]:
    t, c = C.c_double(0), C.c_longlong(0)
    rc = load().rsvd_kstats_get(name.encode(), C.byref(t), C.byref(c))
    return (t.value, c.value) if rc == 0 else (0.0, 0)


class Batch:
    """A batch of independent float32 matrices packed in one device buffer,
    with packed device outputs for rsvd_dev_fixed_rank_batched."""

    def __init__(self, m, n, k, p: int):
        self.m = np.ascontiguousarray(m, dtype=np.int32)
        self.n = np.ascontiguousarray(n, dtype=np.int32)
        self.k = np.ascontiguousarray(k, dtype=np.int32)
        self.p = int(p)
        self.count = len(self.m)
        mn = self.m.astype(np.int64) * self.n
        self.a_off = np.concatenate([[0], np.cumsum(mn)[:-1]]).astype(np.int64)
        mk = self.m.astype(np.int64) * self.k
        nk = self.n.astype(np.int64) * self.k
        self.u_off = np.concatenate([[0], np.cumsum(mk)[:-1]]).astype(np.int64)
        self.v_off = np.concatenate([[0], np.cumsum(nk)[:-1]]).astype(np.int64)
        self.s_off = np.concatenate([[0], np.cumsum(self.k)[:-1]]).astype(np.int64)
        self.A = DeviceMatrix(int(mn.sum()), 1, np.float32)
        self.U = DeviceMatrix(int(mk.sum()), 1, np.float32)
        self.V = DeviceMatrix(int(nk.sum()), 1, np.float32)
        self.S = DeviceMatrix(int(self.k.sum()), 1, np.float64)

    def fill_test_matrices(self, tau, seed: int) -> None:
        tau = np.ascontiguousarray(tau, dtype=np.float64)
        _check(load().rsvd_dev_batch_test_matrices(
            self.count, self.m.ctypes.data, self.n.ctypes.data, tau.ctypes.data, seed,
            self.A.ptr, self.a_off.ctypes.data), "rsvd_dev_batch_test_matrices")

    def run(self, q: int, seed: int) -> int:
        st = C.c_int(0)
        _check(load().rsvd_dev_fixed_rank_batched(
            self.count, self.m.ctypes.data, self.n.ctypes.data, self.k.ctypes.data, self.p, q,
            seed, self.A.ptr, self.a_off.ctypes.data, self.U.ptr, self.u_off.ctypes.data,
            self.S.ptr, self.s_off.ctypes.data, self.V.ptr, self.v_off.ctypes.data,
            C.byref(st)), "rsvd_dev_fixed_rank_batched")
        return st.value

    def matrix(self, i: int) -> np.ndarray:
        """A_i copied to the host (float32, m x n)."""
        a = self.A.to_host()[:, 0]
        o, m, n = self.a_off[i], self.m[i], self.n[i]
        return a[o:o + m * n].reshape(n, m).T

    def results(self):
        """Per matrix (U_i, S_i, V_i) on the host."""
        U, S, V = self.U.to_host()[:, 0], self.S.to_host()[:, 0], self.V.to_host()[:, 0]
        out = []
        for i in range(self.count):
            m, n, k = int(self.m[i]), int(self.n[i]), int(self.k[i])
            u = U[self.u_off[i]:self.u_off[i] + m * k].reshape(k, m).T
            v = V[self.v_off[i]:self.v_off[i] + n * k].reshape(k, n).T
            out.append((u, S[self.s_off[i]:self.s_off[i] + k].copy(), v))
        return out


def batched_last_config() -> str:
    return load().rsvd_batched_last_config().decode()
