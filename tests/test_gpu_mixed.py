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


@pytest.mark.gpu
@pytest.mark.parametrize("prec,bound", [(2, 1e-5), (1, 2e-3), (3, 0.1)])
def test_fixed_rank_mp_against_fp64(gpu_lib, prec, bound):
    """8192 x 2048, sigma_j = 1/(j+1) for j < 512 plus 1e-6 noise, k = 128, p
    = 16, q = 2: the k singular values against the FP64 plain path on the same
    Omega (relative, value by value), orthonormality of the fp32 factors, and
    the residual ||A - U S V^T||_F against the FP64 one."""
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
