"""GPU kernel checks through the C ABI: Gaussian stream, GEMM, QR, eigen/SVD, norms.

Floating-point tolerances are stated per test; the Gaussian stream must be
bit-identical to the oracle's.
"""
import numpy as np
import pytest

from tests.helpers import lowrank_matrix, rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed,offset,count", [(777, 0, 1), (777, 0, 4099), (1, 5, 1000),
                                               (2**50 + 3, 10**10, 777)])
def test_gaussian_stream_bit_identical(gpu_lib, oracle, seed, offset, count):
    got = gpu_lib.gaussian(seed, offset, count)
    want = oracle.gaussian(seed, offset, count)
    assert np.array_equal(got, want)


def test_initialize_random_matrix_seeded(gpu_lib, oracle):
    gpu_lib.set_seed(99)
    try:
        M = gpu_lib.initialize_random_matrix(37, 11)
    finally:
        gpu_lib.use_time_seed()
    want = oracle.gaussian(99, 0, 37 * 11).reshape(11, 37).T
    assert np.array_equal(M, want)


GEMM_SHAPES = [(1, 1, 1), (17, 33, 5), (130, 70, 300), (2000, 110, 1000), (64, 64, 64),
               (129, 65, 17), (300, 7, 5000)]


@pytest.mark.parametrize("m,n,k", GEMM_SHAPES)
def test_gemm_nn(gpu_lib, m, n, k):
    rng = np.random.default_rng(m * 7 + n)
    A, B = rng.standard_normal((m, k)), rng.standard_normal((k, n))
    C = gpu_lib.matrix_matrix_mult(A, B)
    ref = A @ B
    # FP64: error bound ~ k * eps * |A||B|
    assert np.max(np.abs(C - ref)) <= 4 * k * 2.3e-16 * np.max(np.abs(A) @ np.abs(B))


@pytest.mark.parametrize("m,n,k", GEMM_SHAPES)
def test_gemm_tn(gpu_lib, m, n, k):
    rng = np.random.default_rng(m * 5 + k)
    A, B = rng.standard_normal((k, m)), rng.standard_normal((k, n))
    C = gpu_lib.matrix_transpose_matrix_mult(A, B)
    ref = A.T @ B
    assert np.max(np.abs(C - ref)) <= 4 * k * 2.3e-16 * np.max(np.abs(A.T) @ np.abs(B))


@pytest.mark.parametrize("m,n,k", GEMM_SHAPES)
def test_gemm_nt(gpu_lib, m, n, k):
    rng = np.random.default_rng(n * 3 + k)
    A, B = rng.standard_normal((m, k)), rng.standard_normal((n, k))
    C = gpu_lib.matrix_matrix_transpose_mult(A, B)
    ref = A @ B.T
    assert np.max(np.abs(C - ref)) <= 4 * k * 2.3e-16 * np.max(np.abs(A) @ np.abs(B.T))


@pytest.mark.parametrize("ta", ["N", "T"])
def test_fp64_vendor_gemm_switch(gpu_lib, ta):
    """rsvd_set_fp64_vendor_gemm(1) sends a pass-sized FP64 product (8192 x 544
    x 4096) to rocBLAS; both backends meet the FP64 error bound against
    NumPy's product and agree with each other within it."""
    rp = gpu_lib
    lib = rp.load()
    rng = np.random.default_rng(3 if ta == "N" else 4)
    M, N, K = 8192, 544, 4096
    A = rng.standard_normal((M, K) if ta == "N" else (K, M))
    B = rng.standard_normal((K, N))
    dA, dB = rp.DeviceMatrix.from_host(A), rp.DeviceMatrix.from_host(B)
    outs = []
    for vendor in (0, 1):
        dC = rp.DeviceMatrix(M, N)
        lib.rsvd_set_fp64_vendor_gemm(vendor)
        try:
            rp.dev_gemm(ta, "N", dA, dB, dC)
        finally:
            lib.rsvd_set_fp64_vendor_gemm(0)
        outs.append(dC.to_host())
    ref = (A if ta == "N" else A.T) @ B
    bound = 4 * K * 2.3e-16 * np.max(np.abs(A if ta == "N" else A.T) @ np.abs(B))
    for C in outs:
        assert np.max(np.abs(C - ref)) <= bound
    assert np.max(np.abs(outs[0] - outs[1])) <= bound


def test_gemm_shape_mismatch_reports_error(gpu_lib):
    lib = gpu_lib.load()
    A, B, Cm = gpu_lib.to_mat(np.ones((3, 4))), gpu_lib.to_mat(np.ones((4, 5))), gpu_lib.empty_mat(3, 4)
    lib.matrix_matrix_mult(A, B, Cm)
    assert "output is 3x4" in gpu_lib.last_error()
    for x in (A, B, Cm):
        lib.matrix_delete(x)


@pytest.mark.parametrize("m,n", [(50, 3), (2000, 110), (1000, 110), (4096, 256), (300, 300)])
def test_qr_getQ(gpu_lib, m, n):
    rng = np.random.default_rng(m + n)
    M = rng.standard_normal((m, n)) * np.logspace(0, -3, n)
    Q = gpu_lib.QR_factorization_getQ(M)
    assert np.abs(Q.T @ Q - np.eye(n)).max() < 1e-13
    # same range as LAPACK's Q: Q Q^T M = M
    assert np.linalg.norm(Q @ (Q.T @ M) - M) < 1e-13 * np.linalg.norm(M) * np.sqrt(n)


def test_compact_qr_R_upper_and_product(gpu_lib):
    rng = np.random.default_rng(4)
    M = rng.standard_normal((500, 60))
    Q, R = gpu_lib.compact_QR_factorization(M)
    assert np.allclose(np.tril(R, -1), 0)
    assert np.linalg.norm(Q @ R - M) < 1e-13 * np.linalg.norm(M)
    # R agrees with LAPACK's up to row signs
    Rl = np.linalg.qr(M, mode="r")
    assert np.allclose(np.abs(R), np.abs(Rl), rtol=1e-11, atol=1e-12)


def test_qr_householder_fallback_on_rank_deficient(gpu_lib):
    """A repeated column breaks CholeskyQR; the Householder path must take over."""
    rng = np.random.default_rng(5)
    M = rng.standard_normal((400, 30))
    M[:, 7] = M[:, 3]  # exactly dependent columns
    Q = gpu_lib.QR_factorization_getQ(M)
    assert np.abs(Q.T @ Q - np.eye(30)).max() < 1e-13
    assert np.linalg.norm(Q @ (Q.T @ M) - M) < 1e-12 * np.linalg.norm(M)


def _householder(rp, M, want_r=True):
    m, n = M.shape
    Y = rp.DeviceMatrix.from_host(M)
    R = rp.DeviceMatrix(max(n, 1), max(n, 1)) if want_r else None
    rc = rp.load().rsvd_dev_householder_qr(Y.ptr, m, n, m, R.ptr if R else None)
    assert rc == 0, rp.last_error()
    return Y.to_host(), (R.to_host()[:n, :n] if R else None)


@pytest.mark.parametrize("shape", [(400, 30), (300, 100), (1000, 150), (97, 97), (64, 33),
                                   (5, 1), (3000, 64), (70000, 40)])
def test_householder_qr_matches_lapack(gpu_lib, shape):
    """rsvd_dev_householder_qr is dgeqrf + dorgqr: Q and R equal LAPACK's
    (numpy.linalg.qr) to rounding, signs included, for one panel, several
    panels with a ragged last one, square, a single column, and a tall case
    spread over hundreds of workgroups."""
    m, n = shape
    rng = np.random.default_rng(m + 7 * n)
    M = rng.standard_normal((m, n))
    Q, R = _householder(gpu_lib, M)
    Ql, Rl = np.linalg.qr(M)
    assert np.abs(R - Rl).max() <= 1e-13 * np.abs(Rl).max()
    assert np.abs(Q - Ql).max() <= 1e-13 * np.sqrt(n)
    assert np.abs(Q.T @ Q - np.eye(n)).max() <= 1e-14 * np.sqrt(m)


def test_householder_qr_rank_deficient_and_zero_columns(gpu_lib):
    """Exactly dependent and all-zero columns across panel boundaries: Q stays
    orthonormal and spans M's columns (tau = 0 for a zero column, as dlarfg)."""
    rng = np.random.default_rng(12)
    M = rng.standard_normal((2000, 120))
    M[:, 40] = M[:, 3]
    M[:, 77] = 2.0 * M[:, 70] - M[:, 5]
    M[:, 31] = 0.0
    M[:, 100:] = M[:, :20] @ rng.standard_normal((20, 20))
    Q, R = _householder(gpu_lib, M)
    assert np.abs(Q.T @ Q - np.eye(120)).max() < 1e-13
    assert np.linalg.norm(Q @ R - M) < 1e-13 * np.linalg.norm(M)
    assert np.linalg.norm(Q @ (Q.T @ M) - M) < 1e-13 * np.linalg.norm(M)


def test_householder_qr_c2_basis_time(gpu_lib):
    """The fallback at the size the headline's basis has (65536 x 544):
    orthonormal Q, Q R = M, and its time (printed)."""
    import time
    rp = gpu_lib
    lib = rp.load()
    m, n = 65536, 544
    Y = rp.DeviceMatrix(m, n)
    assert lib.rsvd_dev_gaussian(5, m, n, Y.ptr, m) == 0, rp.last_error()
    M = Y.to_host()
    R = rp.DeviceMatrix(n, n)
    times = []
    for _ in range(2):
        Y2 = rp.DeviceMatrix.from_host(M)
        lib.rsvd_sync()
        t0 = time.perf_counter()
        assert lib.rsvd_dev_householder_qr(Y2.ptr, m, n, m, R.ptr) == 0, rp.last_error()
        times.append(time.perf_counter() - t0)
    Q, Rh = Y2.to_host(), R.to_host()
    print(f"householder QR {m}x{n}: {1e3 * min(times):.1f} ms")
    assert np.abs(Q.T @ Q - np.eye(n)).max() < 1e-13
    assert np.linalg.norm(Q @ Rh - M) < 1e-13 * np.linalg.norm(M)


def test_qr_ill_conditioned(gpu_lib):
    rng = np.random.default_rng(6)
    U, _ = np.linalg.qr(rng.standard_normal((1000, 40)))
    V, _ = np.linalg.qr(rng.standard_normal((40, 40)))
    M = (U * np.logspace(0, -12, 40)) @ V.T
    Q = gpu_lib.QR_factorization_getQ(M)
    assert np.abs(Q.T @ Q - np.eye(40)).max() < 1e-12


@pytest.mark.parametrize("n", [2, 5, 33, 110, 300])
def test_symmetric_eig(gpu_lib, n):
    rng = np.random.default_rng(n)
    X = rng.standard_normal((n, n))
    S = X + X.T
    vals, vecs = gpu_lib.compute_evals_and_evecs_of_symm_matrix(S)
    w, W = np.linalg.eigh(S)
    scale = np.abs(w).max()
    assert np.all(np.diff(vals) >= 0)
    assert np.abs(vals - w).max() < 1e-12 * scale
    assert np.linalg.norm(S @ vecs - vecs * vals) < 1e-11 * scale * np.sqrt(n)
    assert np.abs(vecs.T @ vecs - np.eye(n)).max() < 1e-12


def test_symmetric_eig_uses_upper_triangle(gpu_lib):
    rng = np.random.default_rng(9)
    X = rng.standard_normal((20, 20))
    S = X + X.T
    garbage = S.copy()
    garbage[np.tril_indices(20, -1)] = 1e6  # dsyev('U') never reads these
    vals, _ = gpu_lib.compute_evals_and_evecs_of_symm_matrix(garbage)
    assert np.abs(vals - np.linalg.eigvalsh(S)).max() < 1e-12 * np.abs(vals).max()


@pytest.mark.parametrize("n,kind", [(16, "spread"), (17, "spread"), (200, "clustered"),
                                    (544, "clustered"), (300, "rank_deficient")])
def test_dev_sym_eig_psd(gpu_lib, n, kind):
    """The vnum = 2 eigensolver on device: eigenvalues ascending to 1e-12 of the
    largest (LAPACK's own accuracy class), orthonormal eigenvectors, small
    residual; a clustered spectrum and an exactly rank-deficient G (zero
    eigenvalues) must still terminate."""
    import ctypes as C
    from lowrankmatrixdecompositioncodes_amd import rsvdpack as rp
    rng = np.random.default_rng(n)
    Qm, _ = np.linalg.qr(rng.standard_normal((n, n)))
    if kind == "spread":
        lam = np.exp(-np.arange(n) / 3.0)
    elif kind == "clustered":
        lam = np.concatenate([1.0 / (1 + np.arange(30.0)) ** 2,
                              np.full(n - 30, 1e-3) * (1 + 0.2 * rng.random(n - 30))])
    else:
        lam = np.concatenate([rng.random(n // 2) + 0.5, np.zeros(n - n // 2)])
    G = (Qm * lam) @ Qm.T
    G = (G + G.T) / 2
    dG = rp.DeviceMatrix.from_host(G)
    ev, V = rp.DeviceMatrix(n, 1), rp.DeviceMatrix(n, n)
    sw = C.c_int(0)
    lib = rp.load()
    assert lib.rsvd_dev_sym_eig_psd(n, dG.ptr, n, ev.ptr, V.ptr, n, C.byref(sw)) == 0, \
        rp.last_error()
    w = ev.to_host()[:, 0]
    X = V.to_host()
    ref = np.linalg.eigvalsh(G)
    assert np.all(np.diff(w) >= 0)
    assert np.abs(w - ref).max() < 1e-12 * ref.max()
    assert np.abs(X.T @ X - np.eye(n)).max() < 1e-12
    assert np.linalg.norm(G @ X - X * w) < 1e-12 * ref.max() * np.sqrt(n)
    assert 0 < sw.value < 60


@pytest.mark.parametrize("m,n", [(110, 110), (60, 40), (40, 60), (544, 544)])
def test_svd(gpu_lib, m, n):
    rng = np.random.default_rng(m * n)
    M = rng.standard_normal((m, n))
    U, S, Vt = gpu_lib.singular_value_decomposition(M)
    s = np.linalg.svd(M, compute_uv=False)
    k = min(m, n)
    assert rel(np.diag(S)[:k], s) < 1e-12
    assert np.linalg.norm(U @ S @ Vt - M) < 1e-12 * np.linalg.norm(M) * np.sqrt(k)


def test_norms(gpu_lib):
    rng = np.random.default_rng(1)
    A, B = rng.standard_normal((300, 77)), rng.standard_normal((300, 77))
    assert abs(gpu_lib.get_matrix_frobenius_norm(A) - np.linalg.norm(A)) < 1e-13 * np.linalg.norm(A)
    pe = gpu_lib.get_percent_error_between_two_mats(A, B)
    assert abs(pe - 100 * np.linalg.norm(A - B) / np.linalg.norm(A)) < 1e-11


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("m,n,k", [(33, 47, 29), (256, 130, 600), (1000, 64, 4096)])
def test_dev_gemm_all_ops(gpu_lib, ta, tb, dtype, m, n, k):
    rng = np.random.default_rng(m + 3 * n + 7 * k)
    A = rng.standard_normal((k, m) if ta == "T" else (m, k)).astype(dtype)
    B = rng.standard_normal((n, k) if tb == "T" else (k, n)).astype(dtype)
    C0 = rng.standard_normal((m, n)).astype(dtype)
    dA, dB, dC = (gpu_lib.DeviceMatrix.from_host(x) for x in (A, B, C0))
    gpu_lib.dev_gemm(ta, tb, dA, dB, dC, alpha=0.5, beta=-2.0)
    got = dC.to_host()
    opA = A.T if ta == "T" else A
    opB = B.T if tb == "T" else B
    want = 0.5 * (opA.astype(np.float64) @ opB.astype(np.float64)) - 2.0 * C0.astype(np.float64)
    eps = np.finfo(dtype).eps
    bound = 4 * k * eps * (np.abs(opA) @ np.abs(opB)).max() + 4 * eps * np.abs(C0).max()
    assert np.abs(got - want).max() <= bound


def bf16_round(x: np.ndarray) -> np.ndarray:
    """float32 -> nearest bf16 (ties to even), returned as float64."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
@pytest.mark.parametrize("m,n,k", [(33, 47, 29), (300, 130, 700), (1000, 544, 2048)])
def test_gemm_bf16_matches_emulation(gpu_lib, ta, tb, m, n, k):
    """bf16 products with fp32 accumulation: equal to the exact product of the
    bf16-rounded operands up to fp32 accumulation error."""
    rng = np.random.default_rng(m + n + k)
    A = rng.standard_normal((k, m) if ta == "T" else (m, k)).astype(np.float32)
    B = rng.standard_normal((n, k) if tb == "T" else (k, n)).astype(np.float32)
    C0 = rng.standard_normal((m, n)).astype(np.float32)
    dA, dB, dC = (gpu_lib.DeviceMatrix.from_host(x) for x in (A, B, C0))
    gpu_lib.dev_gemm(ta, tb, dA, dB, dC, alpha=1.0, beta=0.5, mode="bf16")
    got = dC.to_host().astype(np.float64)
    opA = bf16_round(A.T if ta == "T" else A)
    opB = bf16_round(B.T if tb == "T" else B)
    want = opA @ opB + 0.5 * C0.astype(np.float64)
    bound = 2 * k * 2.0**-24 * (np.abs(opA) @ np.abs(opB)).max() + 1e-6 * np.abs(C0).max()
    assert np.abs(got - want).max() <= bound


@pytest.mark.parametrize("ta,tb", [("T", "N"), ("N", "N")])
def test_gemm_bf16x3_is_fp32_accurate(gpu_lib, ta, tb):
    rng = np.random.default_rng(11)
    m, n, k = 520, 96, 4096
    A = rng.standard_normal((k, m) if ta == "T" else (m, k)).astype(np.float32)
    B = rng.standard_normal((n, k) if tb == "T" else (k, n)).astype(np.float32)
    dA, dB = gpu_lib.DeviceMatrix.from_host(A), gpu_lib.DeviceMatrix.from_host(B)
    dC = gpu_lib.DeviceMatrix(m, n, np.float32)
    gpu_lib.dev_gemm(ta, tb, dA, dB, dC, mode="bf16x3")
    got = dC.to_host().astype(np.float64)
    opA = (A.T if ta == "T" else A).astype(np.float64)
    opB = (B.T if tb == "T" else B).astype(np.float64)
    want = opA @ opB
    scale = (np.abs(opA) @ np.abs(opB)).max()
    err = np.abs(got - want).max() / scale
    # split-bf16 carries ~16 bits per operand; fp32 accumulation over k
    assert err < 2e-5, err
    # and clearly better than plain bf16
    dC2 = gpu_lib.DeviceMatrix(m, n, np.float32)
    gpu_lib.dev_gemm(ta, tb, dA, dB, dC2, mode="bf16")
    err1 = np.abs(dC2.to_host() - want).max() / scale
    assert err < err1 / 20


@pytest.mark.parametrize("ta", ["N", "T"])
@pytest.mark.parametrize("mode", ["bf16", "bf16x3"])
@pytest.mark.parametrize("M,N,K", [(16400, 544, 300), (16384, 40, 1000), (16448, 600, 77),
                                   (50000, 300, 1000), (49999, 40, 1001), (16384, 544, 4096),
                                   (8192, 64, 16384), (20000, 400, 1001)])
def test_gemm_bf16_streaming_path(gpu_lib, ta, mode, M, N, K):
    """Shapes with a skinny right operand take the A-streaming kernels: the
    64 x 576 tile (A read once, more than one column block when N > 576,
    unaligned rows when K % 4 != 0) for short grids, and with at least a CU's
    worth of row blocks the LDS-ring kernel: 256 x 288 tiles for op(A) = A
    (two column blocks for N = 300, a K % 32 remainder for K = 1001, the k
    range split over eight slices for 8192 x 64 x 16384), 128 x 288 for
    op(A) = A^T, or for op(A) = A^T with 288 < N <= 544 one 128 x 544 tile
    with two LDS stages (16384 x 544 x 4096: two k slices; 50000 x 300 x 1000:
    three slices and the K remainder; 20000 x 400 x 1001: rows of A not
    16-byte aligned).
    bf16: equal to the exact product of the bf16-rounded operands up to fp32
    accumulation; bf16x3: fp32-accurate."""
    rng = np.random.default_rng(M + N + K)
    A = rng.standard_normal((K, M) if ta == "T" else (M, K)).astype(np.float32)
    B = rng.standard_normal((K, N)).astype(np.float32)
    C0 = rng.standard_normal((M, N)).astype(np.float32)
    dA, dB, dC = (gpu_lib.DeviceMatrix.from_host(x) for x in (A, B, C0))
    gpu_lib.dev_gemm(ta, "N", dA, dB, dC, alpha=0.75, beta=0.5, mode=mode)
    got = dC.to_host().astype(np.float64)
    opA = A.T if ta == "T" else A
    if mode == "bf16":
        ea, eb = bf16_round(opA), bf16_round(B)
        want = 0.75 * (ea @ eb) + 0.5 * C0.astype(np.float64)
        bound = 2 * K * 2.0**-24 * (np.abs(ea) @ np.abs(eb)).max() + 1e-6 * np.abs(C0).max()
    else:
        ea, eb = opA.astype(np.float64), B.astype(np.float64)
        want = 0.75 * (ea @ eb) + 0.5 * C0.astype(np.float64)
        bound = 8 * K * 2.0**-24 * (np.abs(ea) @ np.abs(eb)).max() + 1e-6 * np.abs(C0).max()
    assert np.abs(got - want).max() <= bound
