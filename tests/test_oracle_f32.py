"""The FP32 second oracle (the reference's oneAPI single-precision variant).

oracle/cpu_ref.c restates, in float with sgemm / sgeqrf+sorgqr / ssyev /
sgesvd, oneapi_code/rank_revealing_algorithms_one_api.c:72-240
(low_rank_svd_rand_decomp_fixed_rank), :244-304 (low_rank_svd_rand_decomp_fromQB)
and :605-813 (randQB_pb2).  Pinned here against:
* the FP64 oracle on the same Omega (itself pinned in test_oracle_decomp.py):
  the float algorithm is the double one in another precision, so singular
  values, subspaces, the TOL stopping step and the reported rank must agree
  to single-precision accuracy -- a loop count, ordering or truncation slip
  does not;
* closed forms: Q_p^T Q_p = I per block, the TOL norm equals ||M - Q B||, U S V^T = Q B
  for fromQB with the singular values of Q B in descending order;
* the float norm's summation order, on inputs where float and double sums
  differ by construction.
"""
import numpy as np
import pytest

from tests.helpers import lowrank_matrix


@pytest.fixture(scope="module")
def ref():
    from oracle import ref as r
    r.load()
    return r


def _slow(m, n, seed):
    return lowrank_matrix(m, n, 1.0 / (1.0 + np.arange(min(m, n))), seed=seed)


@pytest.mark.parametrize("vnum,q", [(2, 1), (2, 3), (1, 2)])
def test_fixed_rank_f32_matches_fp64_oracle(ref, vnum, q):
    m, n, k, p = 500, 240, 30, 10
    A = _slow(m, n, 3)
    Uf, Sf, Vf = ref.fixed_rank_f32(A.astype(np.float32), k, p, vnum, q, 1, seed=21)
    Ud, Sd, Vd = ref.fixed_rank(A.astype(np.float32).astype(np.float64), k, p, vnum, q, 1,
                                seed=21)
    sf, sd = np.diag(Sf).astype(np.float64), np.diag(Sd)
    # same ordering (vnum 2 ascending / last k, vnum 1 descending / first k)
    assert np.all(np.diff(sd) > 0) if vnum == 2 else np.all(np.diff(sd) < 0)
    assert np.abs(sf - sd).max() <= 2e-6 * sd.max()
    # the leading singular vectors agree up to sign (well-separated values)
    idx = [k - 1, k - 2, k - 3] if vnum == 2 else [0, 1, 2]
    for j in idx:
        assert abs(abs(Uf[:, j].astype(np.float64) @ Ud[:, j]) - 1) < 1e-4
        assert abs(abs(Vf[:, j].astype(np.float64) @ Vd[:, j]) - 1) < 1e-4


def test_fixed_rank_f32_power_iterations_count(ref):
    """q controls q - 1 power iterations as in the double version: q = 2 and
    q = 3 give different sigma on a slowly decaying spectrum, each matching
    the FP64 oracle with the same q (and not with the other)."""
    m, n, k, p = 400, 200, 20, 5
    A = _slow(m, n, 9)
    got = {q: np.diag(ref.fixed_rank_f32(A.astype(np.float32), k, p, 2, q, 1, seed=4)[1])
           for q in (2, 3)}
    want = {q: np.diag(ref.fixed_rank(A.astype(np.float32).astype(np.float64), k, p, 2, q, 1,
                                      seed=4)[1]) for q in (2, 3)}
    gap = np.abs(want[2] - want[3]).max()
    assert gap > 1e-4
    for q in (2, 3):
        assert np.abs(got[q] - want[q]).max() < 0.05 * gap


def test_randqb_pb2_tol_mode_matches_fp64(ref):
    """TOL mode: the same stopping step and rank as randQB_pb_new in double on
    the same RN, the same range to float accuracy, each block of Q
    orthonormal, and the reported norm is ||M - Q B|| to float accuracy."""
    m, n = 600, 300
    sig = np.where(np.arange(n) < 40, 1.0, np.exp(-(np.arange(n) - 40) / 6.0))
    A = lowrank_matrix(m, n, sig, seed=17).astype(np.float32)
    A64 = A.astype(np.float64)
    tol = 1e-4 * np.linalg.norm(A64)
    Qf, Bf, frf, nrm = ref.randQB_pb2(A, 16, 0, tol, 1, 1, seed=33)
    Qd, Bd, frd = ref.randQB_pb_new(A64, 16, 0, tol, 1, 1, seed=33)
    assert frf == frd and Qf.shape == (m, frf) and Bf.shape == (frf, n)
    assert frf % 16 == 0 and 40 < frf < 300
    Q = Qf.astype(np.float64)
    # each block is orthonormal to float accuracy; across blocks the float
    # deflation's rounding (in range of the earlier blocks) leaks into later
    # ones between the even-step re-orthogonalisations: orthogonality of the
    # whole Q degrades to ~1e-4 in single precision
    for b in range(0, frf, 16):
        Qb = Q[:, b:b + 16]
        assert np.abs(Qb.T @ Qb - np.eye(16)).max() < 1e-5
    loss = np.abs(Q.T @ Q - np.eye(frf)).max()
    assert 1e-6 < loss < 1e-3
    res = np.linalg.norm(A64 - Q @ Bf.astype(np.float64))
    assert nrm < tol and abs(nrm - res) <= 1e-3 * res + 1e-6 * np.linalg.norm(A64)
    # the same subspace as the double QB, to the precision float leaves: the
    # blocks that capture values far above the float rounding of A (~1e-7
    # ||A||) agree closely, the last ones (sigma ~ 1e-4) to ~1e-3
    P = Q[:, :64] @ Q[:, :64].T - Qd[:, :64] @ Qd[:, :64].T
    assert np.linalg.norm(P, 2) < 1e-4
    P = Q @ Q.T - Qd @ Qd.T
    assert np.linalg.norm(P, 2) < 1e-2


def test_randqb_pb2_rank_mode_and_kstep_clamp(ref):
    m, n = 300, 200
    A = _slow(m, n, 5).astype(np.float32)
    Q, B, fr, _ = ref.randQB_pb2(A, 12, 3, 0.0, 1, 1, seed=2)
    assert fr == 36 and Q.shape == (m, 36) and B.shape == (36, n)
    # kstep > min(m, n) / 2 is clamped to min(m, n) / 10 (= 20)
    Q, B, fr, _ = ref.randQB_pb2(A, 150, 2, 0.0, 0, 1, seed=2)
    assert fr == 40 and Q.shape == (m, 40)
    # q = 0: the first block spans M RN_0 (RN_0 = first 20 columns of RN)
    rn = ref.gaussian(2, 0, n * 40).reshape(40, n).T
    Y = A.astype(np.float64) @ rn[:, :20]
    Qy, _ = np.linalg.qr(Y)
    Q0 = Q[:, :20].astype(np.float64)
    assert np.linalg.norm(Q0 - Qy @ (Qy.T @ Q0)) < 1e-4


def test_svd_fromqb_f32(ref):
    m, n = 500, 260
    A = _slow(m, n, 12).astype(np.float32)
    Q, B, fr, _ = ref.randQB_pb2(A, 20, 3, 0.0, 1, 1, seed=8)
    U, S, V = ref.svd_fromQB_f32(Q, B)
    assert U.shape == (m, fr) and S.shape == (fr, fr) and V.shape == (n, fr)
    s = np.diag(S).astype(np.float64)
    assert np.all(np.diff(s) <= 0)                       # sgesvd order
    QB = Q.astype(np.float64) @ B.astype(np.float64)
    s_ref = np.linalg.svd(QB, compute_uv=False)[:fr]
    # sigma from the SVD of BB^T: sqrt of float eigenvalues, accurate relative
    # to sigma_1 (squaring halves the relative accuracy of the small ones)
    assert np.abs(s - s_ref).max() <= 1e-5 * s_ref[0]
    top = s > 1e-2 * s[0]
    Uu, Vv = U.astype(np.float64)[:, top], V.astype(np.float64)[:, top]
    assert np.abs(Uu.T @ Uu - np.eye(top.sum())).max() < 1e-5
    assert np.abs(Vv.T @ Vv - np.eye(top.sum())).max() < 1e-4
    recon = (U.astype(np.float64) * s) @ V.astype(np.float64).T
    assert np.linalg.norm(recon - QB) <= 1e-4 * np.linalg.norm(QB)


def test_frobenius_f32_summation_order(ref):
    """Squares summed in float, chunk by chunk: 1.0 followed by many values
    whose squares (1e-8) are below half an ulp of 1 is 1.0 in one sequential
    float sum, but each later chunk sums its own small squares first."""
    N = 4000
    a = np.full((N, 1), 1e-4, dtype=np.float32)
    a[0, 0] = 1.0
    seq = np.float32(0)
    for x in a[:, 0]:
        seq = np.float32(seq + np.float32(x) * np.float32(x))
    assert seq == np.float32(1.0)
    assert ref.frobenius_f32(a, chunks=1) == pytest.approx(float(np.sqrt(seq)), abs=0)
    # 4 chunks of 1000: the last three hold 1000 * 1e-8 = 1e-5 each
    parts = []
    for t in range(4):
        part = np.float32(0)
        for x in a[t * 1000:(t + 1) * 1000, 0]:
            part = np.float32(part + x * x)
        parts.append(part)
    tot = np.float32(0)
    for part in parts:
        tot = np.float32(tot + part)
    got = ref.frobenius_f32(a, chunks=4)
    assert got == pytest.approx(float(np.sqrt(tot)), rel=1e-7)
    assert got > float(np.sqrt(seq)) * (1 + 1e-6)
