"""ORACLE (test infrastructure only): Python access to oracle/cpu_ref.c.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker.  The product (librsvdpack.so) never
loads it.  Builds libcpu_ref.so on demand with the oracle Makefile recipe.
"""
from __future__ import annotations

import ctypes as C
import glob
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "libcpu_ref.so"


class Mat(C.Structure):
    _fields_ = [("nrows", C.c_int), ("ncols", C.c_int), ("d", C.POINTER(C.c_double))]


MatP = C.POINTER(Mat)


class Mat32(C.Structure):
    """The oneAPI variant's mat: int64 dimensions, float data."""
    _fields_ = [("nrows", C.c_int64), ("ncols", C.c_int64), ("d", C.POINTER(C.c_float))]


Mat32P = C.POINTER(Mat32)
_lib = None


def build() -> Path:
    r = subprocess.run(["make", "-s", "-C", str(HERE), "libcpu_ref.so"], capture_output=True,
                       text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    return LIB


def _openblas_path() -> str:
    pats = ["/usr/local/lib/python3*/dist-packages/scipy.libs/libscipy_openblas*.so",
            "/usr/lib/python3*/site-packages/scipy.libs/libscipy_openblas*.so"]
    try:
        import scipy  # noqa: F401
        base = Path(scipy.__file__).resolve().parent.parent / "scipy.libs"
        pats.insert(0, str(base / "libscipy_openblas*.so"))
    except Exception:
        pass
    for p in pats:
        hits = sorted(glob.glob(p))
        if hits:
            return hits[0]
    return ""


def load(nthreads: int | None = None) -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        lib = C.CDLL(str(LIB))
        lib.oref_init.restype = C.c_int
        lib.oref_init.argtypes = [C.c_int, C.c_char_p]
        lib.oref_backend.restype = C.c_char_p
        lib.oref_matrix_new.restype = MatP
        lib.oref_matrix_new.argtypes = [C.c_int, C.c_int]
        lib.oref_matrix_delete.argtypes = [MatP]
        lib.oref_ndtri.restype = C.c_double
        lib.oref_ndtri.argtypes = [C.c_double]
        lib.oref_philox4x64.argtypes = [C.POINTER(C.c_uint64)] * 3
        lib.oref_gaussian.argtypes = [C.c_uint64, C.c_int64, C.c_int64, C.POINTER(C.c_double)]
        lib.oref_set_seed.argtypes = [C.c_uint64]
        lib.oref_set_omega.argtypes = [C.POINTER(C.c_double), C.c_int64]
        lib.oref_fixed_rank.restype = C.c_int
        lib.oref_fixed_rank.argtypes = [MatP, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.POINTER(C.c_int), C.POINTER(MatP), C.POINTER(MatP),
                                        C.POINTER(MatP)]
        lib.oref_randQB_pb.restype = C.c_int
        lib.oref_randQB_pb.argtypes = [MatP, C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.POINTER(MatP), C.POINTER(MatP)]
        lib.oref_randQB_pb_new.restype = C.c_int
        lib.oref_randQB_pb_new.argtypes = [MatP, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int,
                                           C.POINTER(C.c_int), C.POINTER(MatP), C.POINTER(MatP)]
        lib.oref_blockrand.restype = C.c_int
        lib.oref_blockrand.argtypes = [MatP, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int,
                                       C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(MatP),
                                       C.POINTER(MatP), C.POINTER(MatP), C.POINTER(C.c_int)]
        lib.oref_svd_residual.restype = C.c_double
        lib.oref_svd_residual.argtypes = [MatP, MatP, MatP, MatP]
        lib.oref_wall_seconds.restype = C.c_double
        IP = C.POINTER(C.c_int)
        lib.oref_pivoted_qr.restype = C.c_int
        lib.oref_pivoted_qr.argtypes = [MatP, C.POINTER(MatP), C.POINTER(MatP), IP]
        lib.oref_id_blockrand.restype = C.c_int
        lib.oref_id_blockrand.argtypes = [MatP, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int,
                                          C.c_int, IP, IP, C.POINTER(MatP)]
        lib.oref_id_full.restype = C.c_int
        lib.oref_id_full.argtypes = [MatP, C.c_int, IP, C.POINTER(MatP)]
        lib.oref_id_two_sided_blockrand.restype = C.c_int
        lib.oref_id_two_sided_blockrand.argtypes = [MatP, C.c_int, C.c_int, C.c_double, C.c_int,
                                                    C.c_int, C.c_int, IP, IP, IP,
                                                    C.POINTER(MatP), C.POINTER(MatP)]
        lib.oref_cur_blockrand.restype = C.c_int
        lib.oref_cur_blockrand.argtypes = [MatP, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int,
                                           C.c_int, IP, C.POINTER(MatP), C.POINTER(MatP),
                                           C.POINTER(MatP)]
        lib.oref32_matrix_new.restype = Mat32P
        lib.oref32_matrix_new.argtypes = [C.c_int64, C.c_int64]
        lib.oref32_matrix_delete.argtypes = [Mat32P]
        lib.oref32_set_norm_chunks.argtypes = [C.c_int]
        lib.oref32_frobenius.restype = C.c_float
        lib.oref32_frobenius.argtypes = [C.POINTER(C.c_float), C.c_int64]
        lib.oref32_randQB_pb2.restype = C.c_int
        lib.oref32_randQB_pb2.argtypes = [Mat32P, C.c_int64, C.c_int64, C.c_float, C.c_int64,
                                          C.c_int64, C.POINTER(C.c_int64), C.POINTER(Mat32P),
                                          C.POINTER(Mat32P), C.POINTER(C.c_float)]
        lib.oref32_fixed_rank.restype = C.c_int
        lib.oref32_fixed_rank.argtypes = [Mat32P, C.c_int64, C.c_int64, C.c_int64, C.c_int64,
                                          C.c_int64, C.POINTER(Mat32P), C.POINTER(Mat32P),
                                          C.POINTER(Mat32P)]
        lib.oref32_svd_fromQB.restype = C.c_int
        lib.oref32_svd_fromQB.argtypes = [Mat32P, Mat32P, C.POINTER(Mat32P), C.POINTER(Mat32P),
                                          C.POINTER(Mat32P)]
        nt = nthreads if nthreads is not None else int(os.environ.get("OREF_THREADS", "0"))
        if lib.oref_init(nt, _openblas_path().encode()) == 0:
            raise RuntimeError("oracle: no BLAS/LAPACK library could be bound")
        _lib = lib
    elif nthreads:
        _lib.oref_init(nthreads, b"")
    return _lib


def backend() -> str:
    return load().oref_backend().decode()


def _to_mat(a: np.ndarray) -> MatP:
    lib = load()
    a = np.asarray(a, dtype=np.float64)
    M = lib.oref_matrix_new(a.shape[0], a.shape[1])
    if a.size:
        buf = np.asfortranarray(a)
        C.memmove(M.contents.d, buf.ctypes.data, 8 * a.size)
    return M


def _from_mat(M: MatP) -> np.ndarray:
    lib = load()
    m, n = M.contents.nrows, M.contents.ncols
    out = np.empty((m, n), order="F")
    if m * n:
        C.memmove(out.ctypes.data, M.contents.d, 8 * m * n)
    lib.oref_matrix_delete(M)
    return np.ascontiguousarray(out)


class _Omega:
    """Context: Omega from `omega` (array) if given, else the seeded stream."""

    def __init__(self, seed: int, omega: np.ndarray | None):
        self.seed, self.omega = seed, omega

    def __enter__(self):
        lib = load()
        lib.oref_set_seed(self.seed)
        if self.omega is not None:
            self.buf = np.asfortranarray(np.asarray(self.omega, dtype=np.float64)).ravel(order="F")
            lib.oref_set_omega(self.buf.ctypes.data_as(C.POINTER(C.c_double)), self.buf.size)
        return self

    def __exit__(self, *a):
        load().oref_set_omega(None, 0)


def philox4x64(ctr, key):
    lib = load()
    c = (C.c_uint64 * 4)(*ctr)
    k = (C.c_uint64 * 2)(*key)
    o = (C.c_uint64 * 4)()
    lib.oref_philox4x64(c, k, o)
    return [int(o[i]) for i in range(4)]


def ndtri(p: float) -> float:
    return load().oref_ndtri(p)


def gaussian(seed: int, offset: int, count: int) -> np.ndarray:
    out = np.empty(count)
    load().oref_gaussian(seed, offset, count, out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def fixed_rank(M, k, p, vnum, q, s, seed=777, omega=None):
    """Oracle of low_rank_svd_rand_decomp_fixed_rank; returns (U, S, V)."""
    lib = load()
    Mm = _to_mat(M)
    U, S, V = MatP(), MatP(), MatP()
    fr = C.c_int(-1)
    with _Omega(seed, omega):
        rc = lib.oref_fixed_rank(Mm, k, p, vnum, q, s, C.byref(fr), C.byref(U), C.byref(S),
                                 C.byref(V))
    lib.oref_matrix_delete(Mm)
    if rc:
        raise ValueError(f"oref_fixed_rank rc={rc}")
    return _from_mat(U), _from_mat(S), _from_mat(V)


def randQB_pb(M, kstep, nstep, q, s, seed=777, omega=None):
    lib = load()
    Mm = _to_mat(M)
    Q, B = MatP(), MatP()
    with _Omega(seed, omega):
        rc = lib.oref_randQB_pb(Mm, kstep, nstep, q, s, C.byref(Q), C.byref(B))
    lib.oref_matrix_delete(Mm)
    if rc:
        raise ValueError(f"oref_randQB_pb rc={rc}")
    return _from_mat(Q), _from_mat(B)


def randQB_pb_new(M, kstep, nstep, TOL, q, s, seed=777, omega=None):
    lib = load()
    Mm = _to_mat(M)
    Q, B = MatP(), MatP()
    fr = C.c_int(0)
    with _Omega(seed, omega):
        rc = lib.oref_randQB_pb_new(Mm, kstep, nstep, TOL, q, s, C.byref(fr), C.byref(Q),
                                    C.byref(B))
    lib.oref_matrix_delete(Mm)
    if rc:
        raise ValueError(f"oref_randQB_pb_new rc={rc}")
    return _from_mat(Q), _from_mat(B), fr.value


def blockrand(M, k, p, TOL, vnum, kstep, q, s, seed=777, omega=None):
    """Oracle of the blockrand SVD; returns (U, S, V, frank, l)."""
    lib = load()
    Mm = _to_mat(M)
    U, S, V = MatP(), MatP(), MatP()
    fr, l = C.c_int(0), C.c_int(0)
    with _Omega(seed, omega):
        rc = lib.oref_blockrand(Mm, k, p, TOL, vnum, kstep, q, s, C.byref(fr), C.byref(U),
                                C.byref(S), C.byref(V), C.byref(l))
    lib.oref_matrix_delete(Mm)
    if rc:
        raise ValueError(f"oref_blockrand rc={rc}")
    return _from_mat(U), _from_mat(S), _from_mat(V), fr.value, l.value


def _ints(n):
    return np.zeros(max(n, 1), dtype=np.int32)


def _iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def pivoted_qr(M):
    """Oracle of pivotedQR_mkl: (Q, R, I) with M[:, I] = Q R, I 0-based."""
    lib = load()
    Mm = _to_mat(M)
    Q, R = MatP(), MatP()
    I = _ints(M.shape[1])
    rc = lib.oref_pivoted_qr(Mm, C.byref(Q), C.byref(R), _iptr(I))
    lib.oref_matrix_delete(Mm)
    if rc:
        raise ValueError(f"oref_pivoted_qr rc={rc}")
    return _from_mat(Q), _from_mat(R), I[:M.shape[1]].astype(np.int64)


def id_blockrand(M, k, p, TOL, kstep, q, s, seed=777, omega=None):
    """Oracle of id_blockrand_decomp_fixed_rank_or_prec: (I, T, frank)."""
    lib = load()
    Mm = _to_mat(M)
    T = MatP()
    fr = C.c_int(0)
    I = _ints(M.shape[1])
    with _Omega(seed, omega):
        rc = lib.oref_id_blockrand(Mm, k, p, TOL, kstep, q, s, C.byref(fr), _iptr(I), C.byref(T))
    lib.oref_matrix_delete(Mm)
    if rc:
        raise ValueError(f"oref_id_blockrand rc={rc}")
    return I[:M.shape[1]].astype(np.int64), _from_mat(T), fr.value


def id_full(M, k):
    """Oracle of id_decomp_fixed_rank_or_prec for k >= min(m, n): (I, T)."""
    lib = load()
    Mm = _to_mat(M)
    T = MatP()
    I = _ints(M.shape[1])
    rc = lib.oref_id_full(Mm, k, _iptr(I), C.byref(T))
    lib.oref_matrix_delete(Mm)
    if rc:
        raise ValueError(f"oref_id_full rc={rc}")
    return I[:M.shape[1]].astype(np.int64), _from_mat(T)


def id_two_sided_blockrand(M, k, p, TOL, kstep, q, s, seed=777, omega=None):
    """Oracle of id_two_sided_blockrand_decomp_fixed_rank_or_prec:
    (Icol, Irow, T, S, frank)."""
    lib = load()
    Mm = _to_mat(M)
    T, S = MatP(), MatP()
    fr = C.c_int(0)
    Ic, Ir = _ints(M.shape[1]), _ints(M.shape[0])
    with _Omega(seed, omega):
        rc = lib.oref_id_two_sided_blockrand(Mm, k, p, TOL, kstep, q, s, C.byref(fr), _iptr(Ic),
                                             _iptr(Ir), C.byref(T), C.byref(S))
    lib.oref_matrix_delete(Mm)
    if rc:
        raise ValueError(f"oref_id_two_sided_blockrand rc={rc}")
    return (Ic[:M.shape[1]].astype(np.int64), Ir[:M.shape[0]].astype(np.int64), _from_mat(T),
            _from_mat(S), fr.value)


def cur_blockrand(M, k, p, TOL, kstep, q, s, seed=777, omega=None):
    """Oracle of cur_blockrand_decomp_fixed_rank_or_prec: (C, U, R, frank)."""
    lib = load()
    Mm = _to_mat(M)
    Cm, Um, Rm = MatP(), MatP(), MatP()
    fr = C.c_int(0)
    with _Omega(seed, omega):
        rc = lib.oref_cur_blockrand(Mm, k, p, TOL, kstep, q, s, C.byref(fr), C.byref(Cm),
                                    C.byref(Um), C.byref(Rm))
    lib.oref_matrix_delete(Mm)
    if rc:
        raise ValueError(f"oref_cur_blockrand rc={rc}")
    return _from_mat(Cm), _from_mat(Um), _from_mat(Rm), fr.value


def wall() -> float:
    return load().oref_wall_seconds()


# ---- FP32 second oracle: the oneAPI variant (randQB_pb2, fromQB) ----

def _to_mat32(a: np.ndarray) -> Mat32P:
    lib = load()
    a = np.asarray(a, dtype=np.float32)
    M = lib.oref32_matrix_new(a.shape[0], a.shape[1])
    if a.size:
        buf = np.asfortranarray(a)
        C.memmove(M.contents.d, buf.ctypes.data, 4 * a.size)
    return M


def _from_mat32(M: Mat32P) -> np.ndarray:
    lib = load()
    m, n = M.contents.nrows, M.contents.ncols
    out = np.empty((m, n), dtype=np.float32, order="F")
    if m * n:
        C.memmove(out.ctypes.data, M.contents.d, 4 * m * n)
    lib.oref32_matrix_delete(M)
    return np.ascontiguousarray(out)


def frobenius_f32(a: np.ndarray, chunks: int = 16) -> float:
    """The FP32 variant's Frobenius norm: float sums over `chunks` contiguous
    chunks of the column-major data, added in chunk order."""
    lib = load()
    lib.oref32_set_norm_chunks(chunks)
    buf = np.asfortranarray(np.asarray(a, dtype=np.float32)).ravel(order="F")
    return float(lib.oref32_frobenius(buf.ctypes.data_as(C.POINTER(C.c_float)), buf.size))


def randQB_pb2(M, kstep, nstep, TOL, q, s, seed=777, omega=None, norm_chunks=16):
    """FP32 oracle of randQB_pb2 (oneAPI variant); returns (Q, B, frank,
    final_norm) as float32 arrays / Python scalars."""
    lib = load()
    lib.oref32_set_norm_chunks(norm_chunks)
    Mm = _to_mat32(M)
    Q, B = Mat32P(), Mat32P()
    fr = C.c_int64(0)
    nrm = C.c_float(0)
    with _Omega(seed, omega):
        rc = lib.oref32_randQB_pb2(Mm, kstep, nstep, TOL, q, s, C.byref(fr), C.byref(Q),
                                   C.byref(B), C.byref(nrm))
    lib.oref32_matrix_delete(Mm)
    if rc:
        raise ValueError(f"oref32_randQB_pb2 rc={rc}")
    return _from_mat32(Q), _from_mat32(B), fr.value, float(nrm.value)


def svd_fromQB_f32(Q, B):
    """FP32 oracle of low_rank_svd_rand_decomp_fromQB (oneAPI variant):
    returns (U, S, V), S knew x knew diagonal, values descending."""
    lib = load()
    Qm, Bm = _to_mat32(Q), _to_mat32(B)
    U, S, V = Mat32P(), Mat32P(), Mat32P()
    rc = lib.oref32_svd_fromQB(Qm, Bm, C.byref(U), C.byref(S), C.byref(V))
    lib.oref32_matrix_delete(Qm)
    lib.oref32_matrix_delete(Bm)
    if rc:
        raise ValueError(f"oref32_svd_fromQB rc={rc}")
    return _from_mat32(U), _from_mat32(S), _from_mat32(V)


def fixed_rank_f32(M, k, p, vnum, q, s, seed=777, omega=None):
    """FP32 oracle of the oneAPI variant's low_rank_svd_rand_decomp_fixed_rank;
    returns float32 (U, S, V)."""
    lib = load()
    Mm = _to_mat32(M)
    U, S, V = Mat32P(), Mat32P(), Mat32P()
    with _Omega(seed, omega):
        rc = lib.oref32_fixed_rank(Mm, k, p, vnum, q, s, C.byref(U), C.byref(S), C.byref(V))
    lib.oref32_matrix_delete(Mm)
    if rc:
        raise ValueError(f"oref32_fixed_rank rc={rc}")
    return _from_mat32(U), _from_mat32(S), _from_mat32(V)
