"""Pins the oracle's restatement of the reference decompositions (CPU).

The reference ships no golden vectors (Omega is time-seeded) and cannot be
compiled here (its mkl.h / mkl_lapacke.h / mkl_vsl.h are absent from the
image), so each oracle function is pinned by what the mathematics fixes:

* projector identity: rSVD output = top-k SVD of P_K M with
  K = (M M^T)^(q-1) M Omega for the fixed-rank routine (rra.c:101 runs q-1
  power iterations; orthonormalisation does not change the range), computed
  here through numpy without any of the oracle's code;
* exact recovery of a rank-r matrix and k+p = min(m,n) giving the full SVD;
* ordering / truncation conventions (vnum = 2: ascending, last k; vnum = 1:
  descending, first k);
* QB: Q^T Q = I, Q B = Q Q^T M, range(Q) = range(M RN) when q = 0, the
  TOL-mode stopping step and resize, the kstep clamp;
* blockrand's integer arithmetic (p -> kstep, nstep = (k+p)/kstep, frank
  shrink) derived by hand from rra.c:251-279;
* Eckart-Young lower bound.
A plausible slip (q instead of q-1 iterations, a transposed product, reversed
ordering, wrong block count) fails at least one of these.
"""
import numpy as np
import pytest

from oracle import ref
from tests.helpers import haar, lowrank_matrix, projector_svd, rel


def omega(n, l, seed=5):
    return np.random.default_rng(seed).standard_normal((n, l))


def slow_decay(m=300, n=200, seed=3):
    sigma = 1.0 / (1.0 + np.arange(min(m, n))) ** 0.5
    return lowrank_matrix(m, n, sigma, seed=seed), sigma


@pytest.mark.parametrize("q", [0, 1, 2, 3])
@pytest.mark.parametrize("vnum", [1, 2])
def test_fixed_rank_projector_identity(q, vnum):
    A, _ = slow_decay()
    k, p = 20, 8
    Om = omega(200, k + p)
    U, S, V = ref.fixed_rank(A, k, p, vnum, q, 1, omega=Om)
    K = A @ Om
    for _ in range(max(q - 1, 0)):
        K = A @ (A.T @ K)
    Ue, se, Ve = projector_svd(A, K, k)
    s = np.diag(S)
    if vnum == 2:
        assert np.all(np.diff(s) >= 0), "vnum=2 returns ascending singular values"
        s, U, V = s[::-1], U[:, ::-1], V[:, ::-1]
    else:
        assert np.all(np.diff(s) <= 0)
    assert rel(s, se) < 1e-9
    # the rank-k reconstructions agree (sign free)
    assert np.linalg.norm(U * s @ V.T - Ue * se @ Ve.T) < 1e-9 * np.linalg.norm(A)


def test_fixed_rank_loop_count_is_q_minus_1():
    """The projector with q power iterations (MATLAB's count) must NOT match."""
    A, _ = slow_decay()
    k, p, q = 20, 4, 2
    Om = omega(200, k + p)
    _, S, _ = ref.fixed_rank(A, k, p, 1, q, 1, omega=Om)
    K_ref = A @ (A.T @ (A @ Om))
    K_wrong = A @ (A.T @ K_ref)
    _, s_ok, _ = projector_svd(A, K_ref, k)
    _, s_bad, _ = projector_svd(A, K_wrong, k)
    s = np.diag(S)
    assert rel(s, s_ok) < 1e-9
    assert rel(s, s_bad) > 1e-6


@pytest.mark.parametrize("s_param", [1, 2, 3])
def test_fixed_rank_reortho_schedule_keeps_range(s_param):
    A, _ = slow_decay()
    k, p, q = 15, 5, 3
    Om = omega(200, k + p)
    _, S, _ = ref.fixed_rank(A, k, p, 1, q, s_param, omega=Om)
    K = A @ Om
    for _ in range(q - 1):
        K = A @ (A.T @ K)
    _, se, _ = projector_svd(A, K, k)
    assert rel(np.diag(S), se) < 1e-8


@pytest.mark.parametrize("vnum", [1, 2])
def test_exact_low_rank_recovery(vnum):
    sig = np.array([10.0, 5.0, 3.0, 1.0, 0.5, 0.25, 0.1, 0.05])
    A = lowrank_matrix(250, 120, sig, seed=11)
    U, S, V = ref.fixed_rank(A, 8, 6, vnum, 2, 1, omega=omega(120, 14))
    s = np.sort(np.diag(S))[::-1]
    # BB^T squares the conditioning: eps * (sigma_max / sigma_min)^2 ~ 1e-11
    assert rel(s, sig) < (1e-12 if vnum == 1 else 1e-10)
    assert np.linalg.norm(A - U @ S @ V.T) < 1e-12 * np.linalg.norm(A)


def test_full_sample_gives_exact_svd():
    rng = np.random.default_rng(2)
    A = rng.standard_normal((60, 40))
    k, p = 25, 15  # k + p = min(m, n)
    U, S, V = ref.fixed_rank(A, k, p, 1, 1, 1, omega=omega(40, 40))
    s_full = np.linalg.svd(A, compute_uv=False)
    assert rel(np.diag(S), s_full[:k]) < 1e-12
    assert np.allclose(U.T @ U, np.eye(k), atol=1e-13)
    assert np.allclose(V.T @ V, np.eye(k), atol=1e-13)


def test_brute_force_tiny():
    rng = np.random.default_rng(8)
    A = rng.standard_normal((12, 8))
    Om = rng.standard_normal((8, 5))
    U, S, V = ref.fixed_rank(A, 3, 2, 1, 2, 1, omega=Om)
    K = A @ (A.T @ (A @ Om))
    Ue, se, Ve = projector_svd(A, K, 3)
    assert rel(np.diag(S), se) < 1e-12
    assert np.abs(U.T @ Ue).round(10).tolist() == np.eye(3).tolist()


def test_eckart_young_lower_bound():
    A, sigma = slow_decay()
    k = 20
    U, S, V = ref.fixed_rank(A, k, 10, 2, 2, 1, omega=omega(200, 30))
    res = np.linalg.norm(A - U @ S @ V.T)
    assert res >= np.sqrt(np.sum(sigma[k:] ** 2)) * (1 - 1e-12)


def test_vnum1_vnum2_same_spectrum():
    A, _ = slow_decay()
    Om = omega(200, 30)
    _, S1, _ = ref.fixed_rank(A, 20, 10, 1, 2, 1, omega=Om)
    _, S2, _ = ref.fixed_rank(A, 20, 10, 2, 2, 1, omega=Om)
    assert rel(np.diag(S1), np.diag(S2)[::-1]) < 1e-10


def test_seeded_omega_is_the_stream():
    A, _ = slow_decay()
    Om = ref.gaussian(4242, 0, 200 * 30).reshape(30, 200).T  # column-major n x l
    _, S1, _ = ref.fixed_rank(A, 20, 10, 1, 2, 1, seed=4242)
    _, S2, _ = ref.fixed_rank(A, 20, 10, 1, 2, 1, omega=Om)
    assert np.array_equal(S1, S2)


# ---------------------------------------------------------------- QB
@pytest.mark.parametrize("is_new", [0, 1])
def test_qb_q0_range_and_product(is_new):
    A, _ = slow_decay()
    kstep, nstep = 10, 4
    RN = omega(200, kstep * nstep, seed=9)
    if is_new:
        Q, B, fr = ref.randQB_pb_new(A, kstep, nstep, 0.0, 0, 1, omega=RN)
        assert fr == kstep * nstep
    else:
        Q, B = ref.randQB_pb(A, kstep, nstep, 0, 1, omega=RN)
    assert Q.shape == (300, 40) and B.shape == (40, 200)
    assert np.allclose(Q.T @ Q, np.eye(40), atol=1e-12)
    assert np.linalg.norm(Q @ B - Q @ (Q.T @ A)) < 1e-12 * np.linalg.norm(A)
    # range(Q) = range(A RN): projectors agree
    P1 = Q @ Q.T
    Qe, _ = np.linalg.qr(A @ RN)
    assert np.linalg.norm(P1 - Qe @ Qe.T) < 1e-9


@pytest.mark.parametrize("q", [1, 2])
def test_qb_first_block_power_range(q):
    """Step 0 of the QB loop spans (A A^T)^q A RN_0 (q iterations, rra.c:1652)."""
    A, _ = slow_decay()
    kstep = 12
    RN = omega(200, kstep, seed=4)
    Q, B, fr = ref.randQB_pb_new(A, kstep, 1, 0.0, q, 1, omega=RN)
    K = A @ RN
    for _ in range(q):
        K = A @ (A.T @ K)
    Qe, _ = np.linalg.qr(K)
    assert np.linalg.norm(Q @ Q.T - Qe @ Qe.T) < 1e-8


def test_qb_new_tol_mode_stop_and_resize():
    sigma = np.exp(-np.arange(200) / 8.0)
    A = lowrank_matrix(300, 200, sigma, seed=21)
    kstep, TOL = 10, 1e-3
    RN = omega(200, 200, seed=6)
    Q, B, fr = ref.randQB_pb_new(A, kstep, 0, TOL, 1, 1, omega=RN)
    assert fr % kstep == 0 and Q.shape == (300, fr) and B.shape == (fr, 200)
    res = np.linalg.norm(A - Q @ B)
    assert res < TOL
    # one block fewer would not have met TOL
    Qp, Bp = Q[:, : fr - kstep], B[: fr - kstep]
    assert np.linalg.norm(A - Qp @ Bp) >= TOL


def test_qb_new_kstep_clamp():
    A, _ = slow_decay(120, 100)
    # kstep > min(m,n)/2 -> kstep = min(m,n)/10 = 10 (rra.c:1589-1592)
    Q, B, fr = ref.randQB_pb_new(A, 60, 3, 0.0, 0, 1, omega=omega(100, 30))
    assert Q.shape == (120, 30) and fr == 30


def test_qb_pb_reorthogonalises_every_step():
    A, _ = slow_decay()
    Q, B = ref.randQB_pb(A, 10, 6, 1, 1, omega=omega(200, 60, seed=13))
    assert np.abs(Q.T @ Q - np.eye(60)).max() < 1e-13


# ---------------------------------------------------------------- blockrand
def test_blockrand_rank_mode_arithmetic():
    A, _ = slow_decay()
    # k=30, p=5 < kstep=10 and p+kstep < min(m,n): p -> 10, nstep = 40/10 = 4
    U, S, V, fr, l = ref.blockrand(A, 30, 5, 0.0, 1, 10, 1, 1, omega=omega(200, 40))
    assert (fr, l) == (30, 40)
    assert U.shape == (300, 30) and S.shape == (30, 30) and V.shape == (200, 30)


def test_blockrand_integer_nstep_quirk():
    A, _ = slow_decay()
    # k=0, p=0: p -> kstep, nstep = 1 -> one block in rank mode, then the
    # TOL-mode shrink frank = round(20*20/(20+20+1e-6)) = 10
    U, S, V, fr, l = ref.blockrand(A, 0, 0, 1e-3, 2, 20, 1, 1, omega=omega(200, 20))
    assert (fr, l) == (10, 20)
    assert U.shape == (300, 10)


def test_blockrand_tol_mode_reachable_with_negative_k():
    sigma = np.exp(-np.arange(200) / 8.0)
    A = lowrank_matrix(300, 200, sigma, seed=21)
    # k=-1, p=0: p -> 10, nstep = 9/10 = 0 -> randQB_pb_new TOL mode
    U, S, V, fr, l = ref.blockrand(A, -1, 0, 1e-3, 1, 10, 1, 1, omega=omega(200, 200, seed=6))
    _, _, fr_qb = ref.randQB_pb_new(A, 10, 0, 1e-3, 1, 1, omega=omega(200, 200, seed=6))
    assert l == fr_qb
    assert fr == int(np.round(fr_qb / (fr_qb + 10 + 1e-6) * fr_qb))


def test_blockrand_matches_projector_of_qb():
    A, _ = slow_decay()
    RN = omega(200, 40, seed=17)
    U, S, V, fr, l = ref.blockrand(A, 30, 10, 0.0, 1, 10, 1, 1, omega=RN)
    Q, B, _ = ref.randQB_pb_new(A, 10, 4, 0.0, 1, 1, omega=RN)
    Ue, se, Ve = projector_svd(A, Q, 30)
    assert rel(np.diag(S), se) < 1e-10
