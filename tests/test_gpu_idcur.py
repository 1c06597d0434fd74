"""Interpolative and CUR decompositions on the GPU against the oracle.

id_blockrand_decomp_fixed_rank_or_prec, id_two_sided_blockrand_decomp_fixed_rank_or_prec
and cur_blockrand_decomp_fixed_rank_or_prec (rank_revealing_algorithms_intel_mkl.c:1969-2027,
2086-2112, 2262-2332) through the C ABI, on the same seeded RN as the oracle:
* the pivots (column and row index vectors) are the oracle's, exactly;
* T, S and U agree with the oracle's to 1e-9 of their size, C and R are the
  selected columns / rows of M exactly;
* rank and TOL modes, the rank rules of both;
* M is reproduced from the factors (to the accuracy its spectrum allows), and
  the use_*_for_approximation routines print the percent error of that
  product.
"""
import numpy as np
import pytest

from tests.helpers import lowrank_matrix

pytestmark = pytest.mark.gpu


def _case(m=500, n=300, seed=3):
    rng = np.random.default_rng(seed)
    sig = 1.0 / (1.0 + np.arange(min(m, n))) ** 1.5
    A = lowrank_matrix(m, n, sig, seed=seed)
    # distinct column scales keep the pivot choices far from ties
    return A * np.exp(rng.uniform(-0.5, 0.5, n))


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture
def seeded(gpu_lib):
    gpu_lib.set_seed(515)
    yield 515
    gpu_lib.use_time_seed()


@pytest.mark.parametrize("k,tolrel", [(40, 0.0), (-1, 1e-4)])
def test_id_blockrand_matches_oracle(gpu_lib, oracle, seeded, k, tolrel):
    rp = gpu_lib
    A = _case()
    tol = tolrel * np.linalg.norm(A)
    I, T, fr = rp.id_blockrand_decomp_fixed_rank_or_prec(A, k, 10, tol, 16, 1, 1)
    Io, To, fro = oracle.id_blockrand(A, k, 10, tol, 16, 1, 1, seed=seeded)
    assert fr == fro and T.shape == To.shape
    assert np.array_equal(I, Io)
    assert _rel(T, To) <= 1e-9
    err = np.linalg.norm(A[:, I] - A[:, I[:fr]] @ np.hstack([np.eye(fr), T]))
    erro = np.linalg.norm(A[:, Io] - A[:, Io[:fro]] @ np.hstack([np.eye(fro), To]))
    assert abs(err - erro) <= 1e-8 * np.linalg.norm(A)


def test_two_sided_id_and_cur_match_oracle(gpu_lib, oracle, seeded):
    rp = gpu_lib
    A = _case(420, 260, 8)
    Ic, Ir, T, S, fr = rp.id_two_sided_blockrand_decomp_fixed_rank_or_prec(A, 30, 6, 0.0, 12, 1,
                                                                          1)
    Ico, Iro, To, So, fro = oracle.id_two_sided_blockrand(A, 30, 6, 0.0, 12, 1, 1, seed=seeded)
    assert fr == fro == 30
    assert np.array_equal(Ic, Ico) and np.array_equal(Ir, Iro)
    assert _rel(T, To) <= 1e-9 and _rel(S, So) <= 1e-9
    Cm, Um, Rm, frc = rp.cur_blockrand_decomp_fixed_rank_or_prec(A, 30, 6, 0.0, 12, 1, 1)
    Co, Uo, Ro, _ = oracle.cur_blockrand(A, 30, 6, 0.0, 12, 1, 1, seed=seeded)
    assert frc == 30
    assert np.array_equal(Cm, Co) and np.array_equal(Rm, Ro)
    assert _rel(Um, Uo) <= 1e-8
    e = np.linalg.norm(A - Cm @ Um @ Rm)
    eo = np.linalg.norm(A - Co @ Uo @ Ro)
    assert abs(e - eo) <= 1e-8 * np.linalg.norm(A)


def test_exact_rank_reconstruction_and_reports(gpu_lib, seeded, capfd):
    rp = gpu_lib
    lib = rp.load()
    A = lowrank_matrix(600, 350, np.exp(-np.arange(25) / 6.0), seed=12)
    Cm, Um, Rm, fr = rp.cur_blockrand_decomp_fixed_rank_or_prec(A, 25, 5, 0.0, 10, 1, 1)
    assert np.linalg.norm(A - Cm @ Um @ Rm) <= 1e-8 * np.linalg.norm(A)
    I, T, fr = rp.id_blockrand_decomp_fixed_rank_or_prec(A, 25, 5, 0.0, 10, 1, 1)
    MA = A[:, I[:fr]] @ np.hstack([np.eye(fr), T])
    assert np.linalg.norm(A[:, I] - MA) <= 1e-10 * np.linalg.norm(A)
    # the verification routine prints the percent error of that product
    Mm, Tm = rp.to_mat(A), rp.to_mat(T)
    Iv = lib.vector_new(len(I))
    for i, v in enumerate(I):
        lib.vector_set_element(Iv, i, float(v))
    capfd.readouterr()
    lib.use_id_decomp_for_approximation(Mm, Tm, Iv, fr)
    out = capfd.readouterr().out
    pe = 100 * np.linalg.norm(A[:, I] - MA) / np.linalg.norm(A)
    assert "percent error = %f" % pe in out
    Cx, Ux, Rx = rp.to_mat(Cm), rp.to_mat(Um), rp.to_mat(Rm)
    lib.use_cur_decomp_for_approximation(Mm, Cx, Ux, Rx)
    out = capfd.readouterr().out
    pe = 100 * np.linalg.norm(A - Cm @ Um @ Rm) / np.linalg.norm(A)
    assert "percent error = %f" % pe in out
    for x in (Mm, Tm, Cx, Ux, Rx):
        lib.matrix_delete(x)
    lib.vector_delete(Iv)


def test_id_blockrand_larger_tol_mode(gpu_lib, oracle, seeded):
    """2000 x 1200, TOL 1e-6, kstep 64: pivots of the oracle, and a timing line."""
    import time
    rp = gpu_lib
    j = np.arange(600, dtype=np.float64)
    A = lowrank_matrix(2000, 1200, np.where(j < 40, 1.0, np.exp(-(j - 40) / 25.0)), seed=4)
    A *= np.exp(np.random.default_rng(2).uniform(-0.3, 0.3, 1200))
    tol = 1e-6 * np.linalg.norm(A)
    t0 = time.perf_counter()
    I, T, fr = rp.id_blockrand_decomp_fixed_rank_or_prec(A, -1, 10, tol, 64, 1, 1)
    dt = time.perf_counter() - t0
    t0 = time.perf_counter()
    Io, To, fro = oracle.id_blockrand(A, -1, 10, tol, 64, 1, 1, seed=seeded)
    dto = time.perf_counter() - t0
    print(f"id_blockrand 2000x1200 TOL mode: rank {fr}, GPU {dt:.3f} s (host API, with "
          f"transfers), oracle {dto:.3f} s")
    assert fr == fro
    assert np.array_equal(I, Io)
    assert _rel(T, To) <= 1e-8


@pytest.mark.parametrize("shape", [(120, 300), (300, 120), (200, 200), (1, 7), (7, 1)])
def test_pivoted_qr_matches_oracle(gpu_lib, oracle, shape):
    """pivotedQR_mkl (rra.c:924-975): the oracle's pivots (LAPACK dgeqp3) and
    R to 1e-11 of its size, Q orthonormal, M(:, I) = Q R."""
    rp = gpu_lib
    rng = np.random.default_rng(shape[0] * 7 + shape[1])
    M = rng.standard_normal(shape) * np.exp(rng.uniform(-2, 2, shape[1]))
    Q, R, I = rp.pivotedQR_mkl(M)
    Qo, Ro, Io = oracle.pivoted_qr(M)
    assert Q.shape == Qo.shape and R.shape == Ro.shape
    assert np.array_equal(I, Io)
    assert _rel(R, Ro) <= 1e-11
    assert np.abs(Q.T @ Q - np.eye(Q.shape[1])).max() <= 1e-13 * max(shape)
    assert np.abs(Q @ R - M[:, I]).max() <= 1e-12 * np.abs(M).max()
