"""The oracle's interpolative (ID) and CUR decompositions on top of the QB.

oracle/cpu_ref.c restates rank_revealing_algorithms_intel_mkl.c:924-975
(pivotedQR_mkl), :1807-1853 (the full-rank branch of id_decomp_fixed_rank_or_prec),
:1969-2027 (id_blockrand_decomp_fixed_rank_or_prec), :2086-2112 (the
two-sided ID) and :2262-2332 (cur_blockrand_decomp_fixed_rank_or_prec), with
matrix_vector_functions_intel_mkl.c:1477-1494 (dtrsm solve) and :1525-1531
(dgesv).  Pinned against:
* scipy's column-pivoted QR (LAPACK dgeqp3): the same pivots, |R| and Q R =
  M(:, I);
* T = Rk1^-1 Rk2 by scipy's triangular solve;
* exact reconstruction for matrices of exact rank r <= k: the column ID
  M(:, I) = M(:, I(1:k)) [I_k T], the two-sided ID and C U R = M;
* the integer rules: nstep = (k + p) / kstep, the TOL-mode rank shrink
  round(frank / (frank + p + 1e-6) * frank), pivots a permutation.
"""
import numpy as np
import pytest
import scipy.linalg as sl

from tests.helpers import lowrank_matrix


@pytest.fixture(scope="module")
def ref():
    from oracle import ref as r
    r.load()
    return r


@pytest.mark.parametrize("shape", [(40, 90), (90, 40), (64, 64)])
def test_pivoted_qr_matches_dgeqp3(ref, shape):
    rng = np.random.default_rng(sum(shape))
    M = rng.standard_normal(shape) * np.exp(rng.uniform(-3, 3, shape[1]))
    Q, R, I = ref.pivoted_qr(M)
    Q2, R2, P2 = sl.qr(M, pivoting=True, mode="economic")
    assert np.array_equal(I, P2)
    assert np.abs(np.abs(R) - np.abs(R2)).max() <= 1e-12 * np.abs(R2).max()
    assert np.abs(Q @ R - M[:, I]).max() <= 1e-12 * np.abs(M).max()
    d = np.abs(np.diag(R))
    assert np.all(d[:-1] >= d[1:] * (1 - 1e-12))      # rank-revealing order


def _recon_id(M, I, T, k):
    return M[:, I[:k]] @ np.hstack([np.eye(k), T])


def test_id_blockrand_rank_mode_exact_rank(ref):
    r = 18
    A = lowrank_matrix(300, 200, np.exp(-np.arange(r) / 4.0), seed=5)
    I, T, fr = ref.id_blockrand(A, 20, 5, 0.0, 10, 1, 1, seed=11)
    assert fr == 20 and T.shape == (20, 180)
    assert sorted(I.tolist()) == list(range(200))
    assert np.linalg.norm(A[:, I] - _recon_id(A, I, T, 20)) <= 1e-10 * np.linalg.norm(A)


def test_id_blockrand_coefficients_and_nstep(ref):
    """T equals Rk1^-1 Rk2 of the pivoted QR of the QB's B; the QB has
    (k + p) / kstep blocks (integer division): k = 25, p = 9, kstep = 10 ->
    3 blocks, l = 30."""
    A = lowrank_matrix(250, 160, 1.0 / (1.0 + np.arange(160)), seed=6)
    I, T, fr = ref.id_blockrand(A, 25, 9, 0.0, 10, 1, 1, seed=3)
    Q, B = ref.randQB_pb_new(A, 10, 3, 0.0, 1, 1, seed=3)[:2]
    assert B.shape == (30, 160)
    _, R, P = ref.pivoted_qr(B)
    assert np.array_equal(I, P)
    want = sl.solve_triangular(np.triu(R[:25, :25]), R[:25, 25:])
    assert np.abs(T - want).max() <= 1e-10 * max(1.0, np.abs(want).max())


def test_id_blockrand_tol_mode_shrink(ref):
    A = lowrank_matrix(300, 200, np.exp(-np.arange(30) / 3.0), seed=2)
    tol = 1e-6 * np.linalg.norm(A)
    _, _, fq = ref.randQB_pb_new(A, 10, 0, tol, 1, 1, seed=3)
    I, T, fr = ref.id_blockrand(A, -1, 5, tol, 10, 1, 1, seed=3)
    assert fr == int(round(fq / (fq + 5 + 1e-6) * fq))
    assert T.shape == (fr, 200 - fr)


def test_two_sided_id_and_cur_exact_rank(ref):
    r = 15
    A = lowrank_matrix(260, 180, np.exp(-np.arange(r) / 5.0), seed=9)
    Ic, Ir, T, S, fr = ref.id_two_sided_blockrand(A, 16, 4, 0.0, 10, 1, 1, seed=21)
    assert fr == 16 and T.shape == (16, 164) and S.shape == (16, 260 - 16)
    assert sorted(Ir.tolist()) == list(range(260))
    # the row ID is the full ID of MI^T = A(:, Icol(1:k))^T
    MIt = A[:, Ic[:fr]].T
    Ir2, S2 = ref.id_full(MIt, fr)
    assert np.array_equal(Ir, Ir2) and np.allclose(S, S2, rtol=0, atol=1e-12)
    # use_id_two_sided_decomp_for_approximation's product (rra.c:2484-2548)
    Icinv, Irinv = np.argsort(Ic), np.argsort(Ir)
    V = np.vstack([np.eye(fr), T.T])[Icinv]
    U = np.vstack([np.eye(fr), S.T])[Irinv]
    MIJ = A[:, Ic[:fr]][Ir[:fr]]
    assert np.linalg.norm(A - U @ MIJ @ V.T) <= 1e-9 * np.linalg.norm(A)
    Cm, Um, Rm, fr2 = ref.cur_blockrand(A, 16, 4, 0.0, 10, 1, 1, seed=21)
    assert fr2 == 16
    assert np.array_equal(Cm, A[:, Ic[:fr]]) and np.array_equal(Rm, A[Ir[:fr]])
    assert np.linalg.norm(A - Cm @ Um @ Rm) <= 1e-8 * np.linalg.norm(A)
