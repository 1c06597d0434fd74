"""GPU pipeline parity against the oracle (same seeded Omega), through the C ABI.

The plain method is FP64 end to end: singular values and the Frobenius
residual ||A - U S V^T||_F must match the oracle within 1e-10 relative; U and V
match column by column up to sign.  Blocked QB: same block count / frank, same
Q B product and the same range.
"""
import numpy as np
import pytest

from tests.helpers import config1_matrix, lowrank_matrix, rel

pytestmark = pytest.mark.gpu

PARITY = 1e-10


@pytest.fixture(scope="module")
def c1():
    return config1_matrix()


def residual(A, U, S, V):
    return np.linalg.norm(A - U @ S @ V.T)


@pytest.mark.parametrize("vnum", [2, 1])
def test_config1_fixed_rank_parity(gpu_lib, oracle, c1, vnum):
    """2000x1000, sigma_j = exp(-j/20), k=100, p=10, q=2, s=1, shared seeded Omega."""
    A, _ = c1
    seed = 20250101
    gpu_lib.set_seed(seed)
    try:
        U, S, V = gpu_lib.low_rank_svd_rand_decomp_fixed_rank(A, 100, 10, vnum, 2, 1)
    finally:
        gpu_lib.use_time_seed()
    Ur, Sr, Vr = oracle.fixed_rank(A, 100, 10, vnum, 2, 1, seed=seed)
    assert U.shape == (2000, 100) and S.shape == (100, 100) and V.shape == (1000, 100)
    s, sr = np.diag(S), np.diag(Sr)
    assert rel(s, sr) < PARITY
    assert np.count_nonzero(S - np.diag(s)) == 0
    r, rr = residual(A, U, S, V), residual(A, Ur, Sr, Vr)
    assert abs(r - rr) / rr < PARITY
    # singular vectors agree up to sign where the spectrum is well separated
    dots = np.abs(np.sum(U * Ur, axis=0))
    assert np.all(dots > 1 - 1e-8)
    dots = np.abs(np.sum(V * Vr, axis=0))
    assert np.all(dots > 1 - 1e-8)


def test_config1_injected_omega_matches_seeded(gpu_lib, c1):
    A, _ = c1
    seed = 31337
    Om = gpu_lib.gaussian(seed, 0, 1000 * 110).reshape(110, 1000).T
    gpu_lib.set_omega(Om)
    try:
        _, S1, _ = gpu_lib.low_rank_svd_rand_decomp_fixed_rank(A, 100, 10, 2, 2, 1)
    finally:
        gpu_lib.set_omega(None)
    gpu_lib.set_seed(seed)
    try:
        _, S2, _ = gpu_lib.low_rank_svd_rand_decomp_fixed_rank(A, 100, 10, 2, 2, 1)
    finally:
        gpu_lib.use_time_seed()
    assert np.array_equal(S1, S2)


@pytest.mark.parametrize("q,s", [(0, 1), (1, 1), (3, 1), (3, 2), (4, 3)])
@pytest.mark.parametrize("vnum", [1, 2])
def test_fixed_rank_param_grid(gpu_lib, oracle, q, s, vnum):
    sigma = 1.0 / (1.0 + np.arange(200)) ** 0.7
    A = lowrank_matrix(400, 200, sigma, seed=q * 10 + s)
    seed = 1000 + q * 7 + s
    gpu_lib.set_seed(seed)
    try:
        U, S, V = gpu_lib.low_rank_svd_rand_decomp_fixed_rank(A, 25, 7, vnum, q, s)
    finally:
        gpu_lib.use_time_seed()
    Ur, Sr, Vr = oracle.fixed_rank(A, 25, 7, vnum, q, s, seed=seed)
    # power iterations without re-orthogonalisation (s > 1) square the
    # conditioning of the sketch; both sides still agree to 1e-8
    tol = PARITY if s == 1 else 1e-8
    assert rel(np.diag(S), np.diag(Sr)) < tol
    assert abs(residual(A, U, S, V) - residual(A, Ur, Sr, Vr)) < tol * residual(A, Ur, Sr, Vr)


def test_fixed_rank_invalid_arguments(gpu_lib):
    lib = gpu_lib.load()
    import ctypes as C
    M = gpu_lib.to_mat(np.ones((30, 20)))
    U, S, V = gpu_lib.MatP(), gpu_lib.MatP(), gpu_lib.MatP()
    fr = C.c_int(0)
    lib.low_rank_svd_rand_decomp_fixed_rank(M, 15, 10, 2, 2, 1, C.byref(fr), C.byref(U),
                                            C.byref(S), C.byref(V))
    assert not U and not S and not V
    assert "k+p" in gpu_lib.last_error()
    lib.matrix_delete(M)


@pytest.mark.parametrize("is_new,kstep,nstep,q", [(1, 20, 5, 1), (1, 16, 4, 2), (0, 20, 5, 1),
                                                  (0, 10, 3, 0)])
def test_randqb_rank_mode_parity(gpu_lib, oracle, is_new, kstep, nstep, q):
    sigma = np.exp(-np.arange(300) / 30.0)
    A = lowrank_matrix(500, 300, sigma, seed=kstep + nstep)
    seed = 77 + kstep
    gpu_lib.set_seed(seed)
    try:
        if is_new:
            Q, B, fr = gpu_lib.randQB_pb_new(A, kstep, nstep, 0.0, q, 1)
        else:
            Q, B = gpu_lib.randQB_pb(A, kstep, nstep, q, 1)
            fr = Q.shape[1]
    finally:
        gpu_lib.use_time_seed()
    if is_new:
        Qr, Br, frr = oracle.randQB_pb_new(A, kstep, nstep, 0.0, q, 1, seed=seed)
    else:
        Qr, Br = oracle.randQB_pb(A, kstep, nstep, q, 1, seed=seed)
        frr = Qr.shape[1]
    assert fr == frr and Q.shape == Qr.shape and B.shape == Br.shape
    nA = np.linalg.norm(A)
    assert np.linalg.norm(Q @ B - Qr @ Br) < 1e-10 * nA
    assert np.linalg.norm(Q @ Q.T - Qr @ Qr.T) < 1e-8
    assert np.abs(Q.T @ Q - np.eye(Q.shape[1])).max() < 1e-12


def test_randqb_new_tol_mode_parity(gpu_lib, oracle):
    sigma = np.exp(-np.arange(300) / 12.0)
    A = lowrank_matrix(500, 300, sigma, seed=3)
    seed, TOL = 2024, 1e-4
    gpu_lib.set_seed(seed)
    try:
        Q, B, fr = gpu_lib.randQB_pb_new(A, 16, 0, TOL, 1, 1)
    finally:
        gpu_lib.use_time_seed()
    Qr, Br, frr = oracle.randQB_pb_new(A, 16, 0, TOL, 1, 1, seed=seed)
    assert fr == frr and Q.shape == (500, fr) and B.shape == (fr, 300)
    assert np.linalg.norm(A - Q @ B) < TOL
    assert np.linalg.norm(Q @ B - Qr @ Br) < 1e-10 * np.linalg.norm(A)


@pytest.mark.parametrize("k,p,TOL,vnum,kstep", [(60, 10, 0.0, 2, 20), (60, 10, 0.0, 1, 20),
                                                (-1, 0, 1e-3, 2, 16), (-1, 0, 1e-3, 1, 16),
                                                (0, 0, 1e-3, 1, 20)])
def test_blockrand_parity(gpu_lib, oracle, k, p, TOL, vnum, kstep):
    sigma = np.exp(-np.arange(300) / 15.0)
    A = lowrank_matrix(500, 300, sigma, seed=abs(k) + kstep)
    seed = 555
    gpu_lib.set_seed(seed)
    try:
        U, S, V, fr = gpu_lib.low_rank_svd_blockrand_decomp_fixed_rank_or_prec(
            A, k, p, TOL, vnum, kstep, 1, 1)
    finally:
        gpu_lib.use_time_seed()
    Ur, Sr, Vr, frr, l = oracle.blockrand(A, k, p, TOL, vnum, kstep, 1, 1, seed=seed)
    assert fr == frr and U.shape == Ur.shape and S.shape == Sr.shape
    assert rel(np.diag(S), np.diag(Sr)) < 1e-9
    assert abs(residual(A, U, S, V) - residual(A, Ur, Sr, Vr)) < 1e-9 * np.linalg.norm(A)


def test_use_for_approximation_prints(gpu_lib, c1, capfd):
    A, _ = c1
    lib = gpu_lib.load()
    gpu_lib.set_seed(1)
    try:
        U, S, V = gpu_lib.low_rank_svd_rand_decomp_fixed_rank(A, 50, 10, 2, 2, 1)
    finally:
        gpu_lib.use_time_seed()
    Mm, Um, Sm, Vm = (gpu_lib.to_mat(x) for x in (A, U, S, V))
    lib.use_low_rank_svd_for_approximation(Mm, Um, Sm, Vm)
    out = capfd.readouterr().out
    pe = 100 * np.linalg.norm(A - U @ S @ V.T) / np.linalg.norm(A)
    assert "percent error = %f" % pe in out
    for x in (Mm, Um, Sm, Vm):
        lib.matrix_delete(x)
