"""Mixed-precision products and the mixed-precision fixed-rank rSVD on the GPU.

* fp8 GEMM (rsvd_dev_gemm dtype 4): against a NumPy emulation of the same
  quantisation -- OCP e4m3 (3 mantissa bits, bias 7, largest 448, round to
  nearest even), A scaled by 1 / a_c per 128 columns of A (a_c = max |A| /
  448), the skinny operand by 1 / b_n per column, exact products, fp32
  accumulation -- and against the FP64 product with the size of error fp8
  implies.
* rsvd_dev_fixed_rank_mp with each precision against the FP64 plain path on
  the same Omega (the fp32 rounding of the seeded stream): split-bf16 within
  the 1e-5 target, bf16 and fp8 within the error their formats imply.
"""
import ctypes as C

import numpy as np
import pytest

def e4m3(x: np.ndarray) -> np.ndarray:
    """Nearest OCP e4m3 value (ties to even) of each float, as float64."""
    x = np.asarray(x, dtype=np.float64)
    a = np.abs(x)
    e = np.floor(np.log2(np.maximum(a, 2.0**-9)))
    e = np.maximum(e, -6.0)  # subnormals share the step of the smallest binade
    step = 2.0 ** (e - 3)
    q = np.minimum(np.round(a / step) * step, 448.0)
    return np.sign(x) * q


def f32(x):
    return np.asarray(x, dtype=np.float32)


def chunk_scales(A: np.ndarray) -> np.ndarray:
    """a_c = max |A(:, 128c:128c+128)| / 448 in float32 (1 for a zero chunk),
    over A as stored."""
    nc = (A.shape[1] + 127) // 128
    a = np.ones(nc, dtype=np.float32)
    for c in range(nc):
        mx = np.float32(np.abs(A[:, 128 * c:128 * (c + 1)]).max())
        if mx > 0:
            a[c] = mx / np.float32(448.0)
    return a


def emulate_fp8(A: np.ndarray, B: np.ndarray, ta: bool, K32: int):
    """op(A)(:, :K32) @ B(:K32) through the library's fp8 scheme, float32
    arithmetic for the scales as on the device.  Returns the product and
    |op(A)_q| |B_q| (the scale of one product's rounding error)."""
    a = chunk_scales(A)
    opA = (A.T if ta else A)[:, :K32]
    M = opA.shape[0]
    S = f32(B[:K32])
    if ta:   # chunks run along the rows of op(A)
        inv = (np.float32(1.0) / a[np.arange(M) // 128]).astype(np.float32)
        Aq = e4m3(f32(opA) * inv[:, None])
        Sp = S
    else:    # chunks run along K: a_c folded into the skinny operand
        inv = (np.float32(1.0) / a[np.arange(K32) // 128]).astype(np.float32)
        Aq = e4m3(f32(opA) * inv[None, :])
        Sp = (S * a[np.arange(K32) // 128][:, None]).astype(np.float32)
    b = (np.abs(Sp).max(axis=0) / np.float32(448.0)).astype(np.float32)
    b[b == 0] = 1.0
    Sq = e4m3(Sp * (np.float32(1.0) / b)[None, :])
    out = (Aq @ Sq) * b[None, :].astype(np.float64)
    mag = (np.abs(Aq) @ np.abs(Sq)) * b[None, :].astype(np.float64)
    if ta:
        out *= a[np.arange(M) // 128][:, None]
        mag *= a[np.arange(M) // 128][:, None]
    return out, mag


def test_e4m3_emulation_matches_torch():
    """The NumPy e4m3 rounding used below against torch's float8_e4m3fn
    conversion (round to nearest even) over normals, subnormals, ties and the
    saturation edge."""
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.standard_normal(20000) * 10.0 ** rng.uniform(-4, 2.6, 20000),
                        np.arange(-464, 465) / 1.0, np.arange(-64, 65) * 2.0**-10,
                        (np.arange(0, 256) + 0.5) * 2.0**-9])
    x = x[np.abs(x) <= 448.0].astype(np.float32)
    want = torch.from_numpy(x).to(torch.float8_e4m3fn).to(torch.float64).numpy()
    assert np.array_equal(e4m3(x), want)


@pytest.mark.gpu
@pytest.mark.parametrize("ta", ["N", "T"])
@pytest.mark.parametrize("M,N,K", [(1000, 96, 1024), (4096, 40, 1000), (300, 544, 777)])
def test_gemm_fp8_matches_emulation(gpu_lib, ta, M, N, K):
    rng = np.random.default_rng(M + N + K)
    A = rng.standard_normal((K, M) if ta == "T" else (M, K)).astype(np.float32)
    A *= np.exp(rng.uniform(-3, 3, A.shape[1])).astype(np.float32)[None, :]  # uneven columns
    B = rng.standard_normal((K, N)).astype(np.float32)
    dA, dB = gpu_lib.DeviceMatrix.from_host(A), gpu_lib.DeviceMatrix.from_host(B)
    dC = gpu_lib.DeviceMatrix(M, N, np.float32)
    lib = gpu_lib.load()
    rc = lib.rsvd_dev_gemm(4, ta.encode(), b"N", M, N, K, 1.0, dA.ptr, dA.rows, dB.ptr, dB.rows,
                           0.0, dC.ptr, M)
    assert rc == 0, gpu_lib.last_error()
    got = dC.to_host().astype(np.float64)
    opA = A.T if ta == "T" else A
    K32 = K - K % 32
    want, mag = emulate_fp8(A, B, ta == "T", K32)
    # the K % 32 remainder runs in split-bf16 (about fp32)
    want += opA[:, K32:].astype(np.float64) @ B[K32:].astype(np.float64)
    # fp32 accumulation of exact products, plus rare one-step flips of e4m3
    # roundings at a tie of the float32 scaling
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    assert err < 1e-4, err
    assert np.abs(got - want).max() <= 1e-3 * mag.max()
    # and the format's own error against the exact product: a few percent
    exact = opA.astype(np.float64) @ B.astype(np.float64)
    assert np.linalg.norm(got - exact) / np.linalg.norm(exact) < 0.1


def emulate_fp16(A: np.ndarray, B: np.ndarray, ta: bool, K32: int):
    """op(A)(:, :K32) @ B(:K32) through the library's fp16 scheme: a_c =
    max |A chunk| / 32768 per 128 columns of A, A * (1 / a_c) rounded to fp16;
    S' = B (times a_c along k for op(A) = A), b_n = max_k |S'(k, n)| / 32768,
    hi = fp16(S' / b_n), lo = fp16(S' / b_n - hi); exact products hi A + lo A,
    then b_n and (op(A) = A^T) the row's a_c.  float32 arithmetic for the
    scales and conversions, as on the device."""
    f = np.float32
    nc = (A.shape[1] + 127) // 128
    a = np.ones(nc, dtype=f)
    for c in range(nc):
        mx = f(np.abs(A[:, 128 * c:128 * (c + 1)]).max())
        if mx > 0:
            a[c] = mx / f(32768.0)
    inv = (f(1.0) / a).astype(f)
    opA = (A.T if ta else A)[:, :K32]
    M = opA.shape[0]
    if ta:   # chunks along the rows of op(A)
        Aq = (opA * inv[np.arange(M) // 128][:, None]).astype(f).astype(np.float16)
        Sp = B[:K32].astype(f)
    else:    # chunks along K, folded into the skinny operand
        Aq = (opA * inv[np.arange(K32) // 128][None, :]).astype(f).astype(np.float16)
        Sp = (B[:K32].astype(f) * a[np.arange(K32) // 128][:, None]).astype(f)
    b = np.ones(Sp.shape[1], dtype=f)
    mx = np.abs(Sp).max(axis=0).astype(f)
    b[mx > 0] = mx[mx > 0] / f(32768.0)
    v = (Sp * (f(1.0) / b)[None, :]).astype(f)
    hi = v.astype(np.float16)
    lo = (v - hi.astype(f)).astype(np.float16)
    P = Aq.astype(np.float64) @ (hi.astype(np.float64) + lo.astype(np.float64))
    P *= b[None, :].astype(np.float64)
    if ta:
        P *= a[np.arange(M) // 128][:, None].astype(np.float64)
    mag = np.abs(Aq.astype(np.float64)) @ np.abs(v.astype(np.float64)) * b[None, :]
    return P, mag


@pytest.mark.gpu
@pytest.mark.parametrize("ta,M,N,K", [("N", 65536, 288, 4096), ("T", 32768, 288, 8200),
                                      ("N", 1000, 50, 300)])
def test_gemm_fp16_matches_emulation(gpu_lib, ta, M, N, K):
    """rsvd_dev_gemm dtype 6 on the streaming kernel's fp16 path (the first two
    shapes fill the GPU; K = 8200 leaves a remainder of 8 for split-bf16) and
    on the split-bf16 fallback for a small shape."""
    rng = np.random.default_rng(M + N + K)
    A = rng.standard_normal((K, M) if ta == "T" else (M, K)).astype(np.float32)
    A *= np.exp(rng.uniform(-3, 3, A.shape[1])).astype(np.float32)[None, :]
    B = rng.standard_normal((K, N)).astype(np.float32)
    dA, dB = gpu_lib.DeviceMatrix.from_host(A), gpu_lib.DeviceMatrix.from_host(B)
    dC = gpu_lib.DeviceMatrix(M, N, np.float32)
    lib = gpu_lib.load()
    rc = lib.rsvd_dev_gemm(6, ta.encode(), b"N", M, N, K, 1.0, dA.ptr, dA.rows, dB.ptr, dB.rows,
                           0.0, dC.ptr, M)
    assert rc == 0, gpu_lib.last_error()
    got = dC.to_host().astype(np.float64)
    opA = (A.T if ta == "T" else A).astype(np.float64)
    exact = opA @ B.astype(np.float64)
    if M >= 32768:
        K32 = K - K % 32
        want, mag = emulate_fp16(A, B, ta == "T", K32)
        want += opA[:, K32:] @ B[K32:].astype(np.float64)
        # the same quantisation; fp32 accumulation of exact products
        assert np.abs(got - want).max() <= 1e-5 * mag.max()
        # and the format's own error: A's fp16 rounding, 2^-12 relative
        assert np.linalg.norm(got - exact) / np.linalg.norm(exact) < 1e-3
    else:
        assert np.linalg.norm(got - exact) / np.linalg.norm(exact) < 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("prec,bound", [(2, 1e-5), (1, 2e-3), (3, 0.1), (6, 5e-5)])
def test_fixed_rank_mp_against_fp64(gpu_lib, prec, bound):
    """8192 x 2048, sigma_j = 1/(j+1) for j < 512 plus 1e-6 noise, k = 128, p
    = 16, q = 2: the k singular values against the FP64 plain path on the same
    Omega (relative, value by value), orthonormality of the fp32 factors, and
    the residual ||A - U S V^T||_F against the FP64 one.  fp16 passes (6) keep
    A's fp16 rounding (2^-12): ~3e-5 here, where the spectrum has almost no
    noise floor, ~3e-6 on the 65536 x 16384 workload."""
    rp = gpu_lib
    lib = rp.load()
    m, n, r, k, p, q, seed = 8192, 2048, 512, 128, 16, 2, 4242
    sigma = 1.0 / (1.0 + np.arange(r))
    A64 = rp.DeviceMatrix(m, n, np.float64)
    A32 = rp.DeviceMatrix(m, n, np.float32)
    assert lib.rsvd_dev_lowrank_test_matrix(m, n, r, sigma.ctypes.data, 1e-6, 31, A64.ptr, m,
                                            A32.ptr, m) == 0, rp.last_error()
    # the FP64 plain path on the float rounding of the seeded Omega, which is
    # the Omega the mixed path draws for the same seed
    om = rp.gaussian(seed, 0, n * (k + p)).astype(np.float32).astype(np.float64)
    dOm = rp.DeviceMatrix.from_host(om.reshape(k + p, n).T)
    U64, S64, V64 = rp.DeviceMatrix(m, k), rp.DeviceMatrix(k, 1), rp.DeviceMatrix(n, k)
    assert lib.rsvd_dev_fixed_rank(A64.ptr, m, n, m, k, p, 2, q, 1, seed, dOm.ptr, U64.ptr, m,
                                   S64.ptr, V64.ptr, n) == 0, rp.last_error()
    s64 = S64.to_host()[:, 0]
    Um, Sm, Vm = (rp.DeviceMatrix(m, k, np.float32), rp.DeviceMatrix(k, 1),
                  rp.DeviceMatrix(n, k, np.float32))
    assert lib.rsvd_dev_fixed_rank_mp(A32.ptr, m, n, m, k, p, 2, q, 1, seed, None, prec, Um.ptr, m,
                                      Sm.ptr, Vm.ptr, n) == 0, rp.last_error()
    sm = Sm.to_host()[:, 0]
    rel_sigma = np.abs(sm - s64) / s64
    U, V = Um.to_host().astype(np.float64), Vm.to_host().astype(np.float64)
    ortho = max(np.abs(U.T @ U - np.eye(k)).max(), np.abs(V.T @ V - np.eye(k)).max())
    res, res64 = C.c_double(0), C.c_double(0)
    assert lib.rsvd_dev_lowrank_residual(A64.ptr, m, n, m, k, Um.ptr, m, Sm.ptr, Vm.ptr, n, 1,
                                         C.byref(res)) == 0, rp.last_error()
    assert lib.rsvd_dev_lowrank_residual(A64.ptr, m, n, m, k, U64.ptr, m, S64.ptr, V64.ptr, n, 0,
                                         C.byref(res64)) == 0, rp.last_error()
    print(f"precision {prec}: max sigma rel err {rel_sigma.max():.2e}, ortho {ortho:.1e}, "
          f"residual {res.value:.6e} vs FP64 {res64.value:.6e}")
    assert rel_sigma.max() <= bound, (prec, rel_sigma.max())
    # U and V leave the path as float32 after FP64 rotations
    assert ortho <= 5e-5, (prec, ortho)
    assert abs(res.value - res64.value) <= max(bound, 1e-5) * res64.value


@pytest.mark.gpu
def test_fixed_rank_mp_near_degenerate_cluster(gpu_lib):
    """A cluster of 24 nearly equal singular values (sigma = 1 + 1e-6 j, relative
    gaps 1e-6) above a decaying tail, 4000 x 1500, k = 64, p = 16, q = 2.  The
    eigensolver's early stop for the mixed path must not leave rotations
    inside the cluster unfinished (they would show as a loss of orthogonality
    of V = B^T Uhat Sigma^-1 up to the cluster width, ~5e-5): the returned
    sigma agree with the FP64 path to 1e-6 and the cluster's columns of V and U
    are orthonormal to 1e-5, the level the split-bf16 U and V products leave
    (16-bit operand splits, ~4e-6 measured)."""
    rp = gpu_lib
    lib = rp.load()
    m, n, k, p, q, seed = 4000, 1500, 64, 16, 2, 99
    sigma = np.concatenate([1.0 + 1e-6 * np.arange(24)[::-1],
                            0.5 * np.exp(-np.arange(200) / 20.0)])
    A64 = rp.DeviceMatrix(m, n, np.float64)
    A32 = rp.DeviceMatrix(m, n, np.float32)
    assert lib.rsvd_dev_lowrank_test_matrix(m, n, len(sigma), sigma.ctypes.data, 0.0, 5, A64.ptr,
                                            m, A32.ptr, m) == 0, rp.last_error()
    om = rp.gaussian(seed, 0, n * (k + p)).astype(np.float32).astype(np.float64)
    dOm = rp.DeviceMatrix.from_host(om.reshape(k + p, n).T)
    U64, S64, V64 = rp.DeviceMatrix(m, k), rp.DeviceMatrix(k, 1), rp.DeviceMatrix(n, k)
    assert lib.rsvd_dev_fixed_rank(A64.ptr, m, n, m, k, p, 2, q, 1, seed, dOm.ptr, U64.ptr, m,
                                   S64.ptr, V64.ptr, n) == 0, rp.last_error()
    Um, Sm, Vm = (rp.DeviceMatrix(m, k, np.float32), rp.DeviceMatrix(k, 1),
                  rp.DeviceMatrix(n, k, np.float32))
    assert lib.rsvd_dev_fixed_rank_mp(A32.ptr, m, n, m, k, p, 2, q, 1, seed, None, 2, Um.ptr, m,
                                      Sm.ptr, Vm.ptr, n) == 0, rp.last_error()
    s64, sm = S64.to_host()[:, 0], Sm.to_host()[:, 0]
    assert np.abs(sm - s64).max() <= 1e-6 * s64.max()
    # vnum = 2 returns ascending values: the cluster is the last 24 columns
    V = Vm.to_host().astype(np.float64)[:, -24:]
    U = Um.to_host().astype(np.float64)[:, -24:]
    ov = np.abs(V.T @ V - np.eye(24)).max()
    ou = np.abs(U.T @ U - np.eye(24)).max()
    print(f"cluster: sigma diff {np.abs(sm - s64).max():.2e}, V^T V - I {ov:.2e}, "
          f"U^T U - I {ou:.2e}")
    assert ov <= 1e-5 and ou <= 1e-5, (ov, ou)


def _blockrand_case(rp, m, n, seed):
    """A = U diag(sigma) V^T with a 50-value plateau at 1, then exp(-(j - 50)/40)
    down to 1e-8 (float32 and its exact double widening)."""
    lib = rp.load()
    j = np.arange(min(n, 800), dtype=np.float64)
    sigma = np.where(j < 50, 1.0, np.exp(-(j - 50) / 40.0))
    sigma = sigma[sigma >= 1e-8]
    A64 = rp.DeviceMatrix(m, n, np.float64)
    A32 = rp.DeviceMatrix(m, n, np.float32)
    assert lib.rsvd_dev_lowrank_test_matrix(m, n, len(sigma), sigma.ctypes.data, 0.0, seed,
                                            A64.ptr, m, A32.ptr, m) == 0, rp.last_error()
    return A64, A32, sigma


def _blockrand(rp, mp, A, m, n, k, p, tol, kstep, q, prec=2):
    lib = rp.load()
    fr = C.c_int(0)
    u, sv, v = C.c_void_p(), C.c_void_p(), C.c_void_p()
    if mp:
        rc = lib.rsvd_dev_blockrand_mp(A.ptr, m, n, m, k, p, tol, 2, kstep, q, 1, 77, None, prec,
                                       C.byref(fr), C.byref(u), C.byref(sv), C.byref(v))
    else:
        rc = lib.rsvd_dev_blockrand(A.ptr, m, n, m, k, p, tol, 2, kstep, q, 1, 77, None,
                                    C.byref(fr), C.byref(u), C.byref(sv), C.byref(v))
    assert rc == 0, rp.last_error()
    f = fr.value
    s = np.empty(f)
    lib.rsvd_memcpy_d2h(s.ctypes.data, sv.value, 8 * f)
    return f, s, u, sv, v


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [2, 1, 3])
def test_blockrand_mp_tol_mode(gpu_lib, prec):
    """TOL mode (k < 0), TOL = 1e-6 ||A||_F, kstep = 64, q = 1 on 6000 x 1500:
    the mixed path's explicit residual ||A - U S V^T||_F (FP64, against the
    double A) meets TOL, its rank is the FP64 path's, and its singular values
    agree with the FP64 path's to fp32 level relative to sigma_1."""
    rp = gpu_lib
    lib = rp.load()
    m, n, kstep, q = 6000, 1500, 64, 1
    A64, A32, sigma = _blockrand_case(rp, m, n, 4321)
    anorm = float(np.sqrt(np.sum(sigma ** 2)))
    tol = 1e-6 * anorm
    f64_, s64, *bufs64 = _blockrand(rp, False, A64, m, n, -1, 10, tol, kstep, q)
    fmp, smp, u, sv, v = _blockrand(rp, True, A32, m, n, -1, 10, tol, kstep, q, prec)
    res = C.c_double(0)
    assert lib.rsvd_dev_lowrank_residual(A64.ptr, m, n, m, fmp, u.value, m, sv.value, v.value, n,
                                         1, C.byref(res)) == 0, rp.last_error()
    res64 = C.c_double(0)
    assert lib.rsvd_dev_lowrank_residual(A64.ptr, m, n, m, f64_, bufs64[0].value, m,
                                         bufs64[1].value, bufs64[2].value, n, 0,
                                         C.byref(res64)) == 0, rp.last_error()
    for b in (u, sv, v, *bufs64):
        lib.rsvd_dev_free(b.value)
    ds = np.abs(np.sort(smp)[::-1][:min(fmp, f64_)] - np.sort(s64)[::-1][:min(fmp, f64_)]).max()
    print(f"precision {prec}: frank {fmp} (FP64 {f64_}), residual {res.value / anorm:.3e} "
          f"(FP64 {res64.value / anorm:.3e}) x ||A||, max |dsigma| {ds:.2e}")
    # split-bf16 sketches find the FP64 path's blocks; bf16 / fp8 sketches
    # capture less of R per block and may need more blocks
    if prec == 2:
        assert fmp == f64_
    else:
        assert fmp >= f64_
    # the TOL test is the QB's (test_randqb_mp_tol_mode); the shrunk rank of
    # rra.c:278 then drops the last ~p columns, so the returned factors sit a
    # little above TOL in either path (FP64 here: 0.90 TOL)
    assert res.value <= 1.5 * tol and res64.value <= 1.5 * tol
    assert ds <= 1e-6 * sigma[0]


@pytest.mark.gpu
@pytest.mark.parametrize("kstep", [40, 96])
def test_blockrand_mp_tol_mode_other_kstep(gpu_lib, kstep):
    """The deflation's two kernels: kstep = 40 gives k blocks of 40 and 80
    columns (not multiples of 32: the register-staged tiled kernel, no
    hi-only range); kstep = 96 gives 96 and 192 (the LDS-ring kernel with an
    odd k-block count and the hi-only / six-product boundary after block 3).
    Same checks as the kstep = 64 case, split-bf16 sketches."""
    rp = gpu_lib
    lib = rp.load()
    m, n, q = 6000, 1500, 1
    A64, A32, sigma = _blockrand_case(rp, m, n, 4321)
    anorm = float(np.sqrt(np.sum(sigma ** 2)))
    tol = 1e-6 * anorm
    f64_, s64, *bufs64 = _blockrand(rp, False, A64, m, n, -1, 10, tol, kstep, q)
    fmp, smp, u, sv, v = _blockrand(rp, True, A32, m, n, -1, 10, tol, kstep, q, 2)
    res = C.c_double(0)
    assert lib.rsvd_dev_lowrank_residual(A64.ptr, m, n, m, fmp, u.value, m, sv.value, v.value, n,
                                         1, C.byref(res)) == 0, rp.last_error()
    for b in (u, sv, v, *bufs64):
        lib.rsvd_dev_free(b.value)
    kk = min(fmp, f64_)
    ds = np.abs(np.sort(smp)[::-1][:kk] - np.sort(s64)[::-1][:kk]).max()
    print(f"kstep {kstep}: frank {fmp} (FP64 {f64_}), residual {res.value / anorm:.3e} x ||A||, "
          f"max |dsigma| {ds:.2e}")
    assert fmp >= f64_
    assert res.value <= 1.5 * tol
    assert ds <= 1e-6 * sigma[0]


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [2, 1])
def test_randqb_mp_tol_mode(gpu_lib, prec):
    """randQB_pb_new in TOL mode on the float32 A (6000 x 1500, kstep = 64,
    q = 1, TOL = 1e-6 ||A||_F): the reported norm is below TOL and equals the
    explicit FP64 ||A - Q B||_F of the returned factors (1%), Q is orthonormal
    to fp32 level, and split-bf16 sketches stop at the FP64 path's block."""
    rp = gpu_lib
    lib = rp.load()
    m, n, kstep, q = 6000, 1500, 64, 1
    A64, A32, sigma = _blockrand_case(rp, m, n, 4321)
    anorm = float(np.sqrt(np.sum(sigma ** 2)))
    tol = 1e-6 * anorm
    lmax = lib.rsvd_qb_max_cols(m, n, kstep, 0, 1)
    Q = rp.DeviceMatrix(m, lmax, np.float32)
    Bt = rp.DeviceMatrix(n, lmax, np.float32)
    fr, fn = C.c_int(0), C.c_double(0)
    assert lib.rsvd_dev_randqb_mp(A32.ptr, m, n, m, kstep, 0, tol, q, 1, 1, 77, None, prec, Q.ptr,
                                  m, Bt.ptr, n, C.byref(fr), C.byref(fn)) == 0, rp.last_error()
    Q64, B64 = rp.DeviceMatrix(m, lmax), rp.DeviceMatrix(lmax, n)
    fr64, fn64 = C.c_int(0), C.c_double(0)
    assert lib.rsvd_dev_randqb(A64.ptr, m, n, m, kstep, 0, tol, q, 1, 1, 77, None, Q64.ptr, m,
                               B64.ptr, lmax, C.byref(fr64), C.byref(fn64)) == 0, rp.last_error()
    f = fr.value
    A = A32.to_host().astype(np.float64)
    Qh = Q.to_host()[:, :f].astype(np.float64)
    Bh = Bt.to_host()[:, :f].astype(np.float64)
    explicit = np.linalg.norm(A - Qh @ Bh.T)
    orth = np.abs(Qh.T @ Qh - np.eye(f)).max()
    print(f"precision {prec}: columns {f} (FP64 {fr64.value}), norm {fn.value / anorm:.4e} "
          f"(FP64 {fn64.value / anorm:.4e}), explicit {explicit / anorm:.4e}, orth {orth:.1e}")
    assert fn.value < tol
    assert abs(explicit - fn.value) <= 0.01 * fn.value
    assert orth <= 1e-6
    if prec == 2:
        assert f == fr64.value


@pytest.mark.gpu
def test_blockrand_mp_rank_mode(gpu_lib):
    """Rank mode (k = 120, p = 10, kstep = 64 -> the p -> kstep rule of
    rra.c:260-262): singular values against the FP64 path on the same seeds."""
    rp = gpu_lib
    lib = rp.load()
    m, n = 4000, 1200
    A64, A32, sigma = _blockrand_case(rp, m, n, 99)
    f64_, s64, *b64 = _blockrand(rp, False, A64, m, n, 120, 10, 0.0, 64, 1)
    fmp, smp, *bmp = _blockrand(rp, True, A32, m, n, 120, 10, 0.0, 64, 1)
    for b in (*b64, *bmp):
        lib.rsvd_dev_free(b.value)
    assert fmp == f64_ == 120
    rel = np.abs(smp - s64) / sigma[0]
    print(f"rank mode: max |dsigma| / sigma_1 {rel.max():.2e}")
    assert rel.max() <= 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("ta,tb", [("N", "T"), ("T", "N"), ("N", "N"), ("T", "T")])
@pytest.mark.parametrize("M,N,K", [(1000, 700, 64), (517, 300, 512), (4096, 256, 256)])
def test_gemm_bf16x6_is_fp32_exact(gpu_lib, ta, tb, M, N, K):
    """rsvd_dev_gemm dtype 5 (three bf16 planes per operand, six products):
    C = C0 - op(A) op(B) with C0 the float rounding of op(A) op(B) cancels to
    the fp32 accumulation error, where the 16-bit split (dtype 3) leaves
    ~2^-16."""
    rng = np.random.default_rng(M + K)
    opA = rng.standard_normal((M, K)).astype(np.float32)
    opB = rng.standard_normal((K, N)).astype(np.float32)
    A = np.ascontiguousarray(opA.T) if ta == "T" else opA
    B = np.ascontiguousarray(opB.T) if tb == "T" else opB
    exact = opA.astype(np.float64) @ opB.astype(np.float64)
    C0 = exact.astype(np.float32)
    mag = np.abs(opA).astype(np.float64) @ np.abs(opB).astype(np.float64)
    lib = gpu_lib.load()
    errs = {}
    for dtype in (5, 3):
        dA, dB = gpu_lib.DeviceMatrix.from_host(A), gpu_lib.DeviceMatrix.from_host(B)
        dC = gpu_lib.DeviceMatrix.from_host(C0)
        assert lib.rsvd_dev_gemm(dtype, ta.encode(), tb.encode(), M, N, K, -1.0, dA.ptr, dA.rows,
                                 dB.ptr, dB.rows, 1.0, dC.ptr, M) == 0, gpu_lib.last_error()
        got = dC.to_host().astype(np.float64)
        want = C0.astype(np.float64) - exact
        errs[dtype] = float((np.abs(got - want) / mag).max())
    print(f"max |err| / (|A||B|^T): x6 {errs[5]:.2e}, x3 {errs[3]:.2e}")
    # fp32 accumulation over K / 32 MFMA steps plus the final subtraction
    assert errs[5] <= (K / 32 + 8) * 2.0**-24
    assert errs[3] > 4 * errs[5]


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [1, 0, 3])
@pytest.mark.parametrize("n", [544, 300])
def test_gram_upper_tiles_mirrored(gpu_lib, dtype, n):
    """A^T A with one pointer for both operands and beta = 0 forms only the
    tiles on and above the diagonal and mirrors the rest: the result is
    exactly symmetric and agrees with the same product taken as a general GEMM
    (B a separate copy of A) to the accumulation-order level of each dtype
    (0 FP64, 1 fp32, 3 split-bf16)."""
    rng = np.random.default_rng(n + dtype)
    m = 65536 if dtype != 0 else 16384
    np_t = np.float64 if dtype == 0 else np.float32
    A = rng.standard_normal((m, n)).astype(np_t)
    lib = gpu_lib.load()
    dA = gpu_lib.DeviceMatrix.from_host(A)
    dB = gpu_lib.DeviceMatrix.from_host(A.copy())
    out = {}
    for name, bptr in (("gram", dA.ptr), ("general", dB.ptr)):
        dC = gpu_lib.DeviceMatrix.from_host(np.full((n, n), np.nan, dtype=np_t))
        assert lib.rsvd_dev_gemm(dtype, b"T", b"N", n, n, m, 1.0, dA.ptr, m, bptr, m, 0.0,
                                 dC.ptr, n) == 0, gpu_lib.last_error()
        out[name] = dC.to_host().astype(np.float64)
    G, H = out["gram"], out["general"]
    assert np.isfinite(G).all()
    assert np.array_equal(G, G.T)
    exact = A.astype(np.float64).T @ A.astype(np.float64)
    scale = np.sqrt(np.outer(np.diag(exact), np.diag(exact)))
    tol = {0: 1e-13, 1: 2e-6, 3: 5e-5}[dtype]
    assert (np.abs(G - H) / scale).max() <= tol
    assert (np.abs(G - exact) / scale).max() <= tol


@pytest.mark.gpu
def test_fixed_rank_mp_ill_conditioned_basis(gpu_lib):
    """A steep spectrum, sigma_j = 10^(-3 j / 600) (the sketch basis Y = A Z
    has condition ~1e3, past what a split-bf16 CholeskyQR pass resolves): the
    final CholeskyQR2 falls back to an extra exact pass or to Householder, and
    the result still matches the FP64 path to fp32 level (relative to sigma_1)
    with orthonormal factors."""
    rp = gpu_lib
    lib = rp.load()
    m, n, r, k, p, q, seed = 8192, 2048, 600, 128, 16, 2, 777
    sigma = 10.0 ** (-3.0 * np.arange(r) / r)
    A64 = rp.DeviceMatrix(m, n, np.float64)
    A32 = rp.DeviceMatrix(m, n, np.float32)
    assert lib.rsvd_dev_lowrank_test_matrix(m, n, r, sigma.ctypes.data, 0.0, 5, A64.ptr, m,
                                            A32.ptr, m) == 0, rp.last_error()
    om = rp.gaussian(seed, 0, n * (k + p)).astype(np.float32).astype(np.float64)
    dOm = rp.DeviceMatrix.from_host(om.reshape(k + p, n).T)
    U64, S64, V64 = rp.DeviceMatrix(m, k), rp.DeviceMatrix(k, 1), rp.DeviceMatrix(n, k)
    assert lib.rsvd_dev_fixed_rank(A64.ptr, m, n, m, k, p, 2, q, 1, seed, dOm.ptr, U64.ptr, m,
                                   S64.ptr, V64.ptr, n) == 0, rp.last_error()
    Um, Sm, Vm = (rp.DeviceMatrix(m, k, np.float32), rp.DeviceMatrix(k, 1),
                  rp.DeviceMatrix(n, k, np.float32))
    assert lib.rsvd_dev_fixed_rank_mp(A32.ptr, m, n, m, k, p, 2, q, 1, seed, None, 2, Um.ptr, m,
                                      Sm.ptr, Vm.ptr, n) == 0, rp.last_error()
    s64, sm = S64.to_host()[:, 0], Sm.to_host()[:, 0]
    U, V = Um.to_host().astype(np.float64), Vm.to_host().astype(np.float64)
    ortho = max(np.abs(U.T @ U - np.eye(k)).max(), np.abs(V.T @ V - np.eye(k)).max())
    err = np.abs(sm - s64).max() / s64.max()
    print(f"steep spectrum: max |dsigma| / sigma_1 {err:.2e}, ortho {ortho:.1e}")
    assert err <= 1e-6
    assert ortho <= 5e-5
