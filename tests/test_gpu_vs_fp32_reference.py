"""The mixed-precision paths against the reference's own single-precision variant.

The FP32 second oracle restates the reference's oneAPI code in float
(oneapi_code/rank_revealing_algorithms_one_api.c:72-240 fixed rank, :605-813
randQB_pb2, :244-304 fromQB).  On the same A and Omega, the error of each
against the FP64 path (reference precision) is measured; the mixed paths must
be at least as accurate as the reference's own float arithmetic:

* fixed rank, 16384 x 4096 float32, sigma_j = 1/(j + 1) (rank 1024) plus 1e-6
  noise, k = 256, p = 16, q = 2, vnum = 2: max relative sigma error and the
  residual's relative difference, split-bf16 path vs the FP32 reference;
* blocked QB, 6000 x 1500, TOL = 1e-6 ||A||_F, kstep = 64, q = 1: the explicit
  FP64 residual ||A - Q B||_F and the orthogonality of Q, mixed QB
  (rsvd_dev_randqb_mp) vs randQB_pb2; the singular values of Q B (fromQB on
  the FP32 side) against the FP64 QB's.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _test_matrix(rp, m, n, sigma, noise, seed):
    lib = rp.load()
    A64 = rp.DeviceMatrix(m, n, np.float64)
    A32 = rp.DeviceMatrix(m, n, np.float32)
    assert lib.rsvd_dev_lowrank_test_matrix(m, n, len(sigma), sigma.ctypes.data, noise, seed,
                                            A64.ptr, m, A32.ptr, m) == 0, rp.last_error()
    return A64, A32


def test_fixed_rank_mixed_vs_fp32_reference(gpu_lib, oracle):
    rp = gpu_lib
    lib = rp.load()
    m, n, k, p, q, seed = 16384, 4096, 256, 16, 2, 606
    l = k + p
    sigma = 1.0 / (1.0 + np.arange(1024, dtype=np.float64))
    A64, A32 = _test_matrix(rp, m, n, sigma, 1e-6, 71)
    om = rp.gaussian(seed, 0, n * l).astype(np.float32).astype(np.float64).reshape(l, n).T
    dOm = rp.DeviceMatrix.from_host(om)
    U64, S64, V64 = rp.DeviceMatrix(m, k), rp.DeviceMatrix(k, 1), rp.DeviceMatrix(n, k)
    assert lib.rsvd_dev_fixed_rank(A64.ptr, m, n, m, k, p, 2, q, 1, seed, dOm.ptr, U64.ptr, m,
                                   S64.ptr, V64.ptr, n) == 0, rp.last_error()
    Um, Sm, Vm = (rp.DeviceMatrix(m, k, np.float32), rp.DeviceMatrix(k, 1),
                  rp.DeviceMatrix(n, k, np.float32))
    assert lib.rsvd_dev_fixed_rank_mp(A32.ptr, m, n, m, k, p, 2, q, 1, seed, None, 2, Um.ptr, m,
                                      Sm.ptr, Vm.ptr, n) == 0, rp.last_error()
    s64, smp = S64.to_host()[:, 0], Sm.to_host()[:, 0]
    Ah = A32.to_host()
    Uf, Sf, Vf = oracle.fixed_rank_f32(Ah, k, p, 2, q, 1, omega=om)
    sf = np.diag(Sf).astype(np.float64)

    def resid(U, S, V, f32):
        out = C.c_double(0)
        assert lib.rsvd_dev_lowrank_residual(A64.ptr, m, n, m, k, U, m, S, V, n, f32,
                                             C.byref(out)) == 0, rp.last_error()
        return out.value

    dUf = rp.DeviceMatrix.from_host(Uf)
    dSf = rp.DeviceMatrix.from_host(sf.reshape(-1, 1))
    dVf = rp.DeviceMatrix.from_host(Vf)
    r64 = resid(U64.ptr, S64.ptr, V64.ptr, 0)
    rmp = resid(Um.ptr, Sm.ptr, Vm.ptr, 1)
    rf = resid(dUf.ptr, dSf.ptr, dVf.ptr, 1)
    err_mp = float(np.max(np.abs(smp - s64) / s64))
    err_f = float(np.max(np.abs(sf - s64) / s64))
    print(f"fixed rank {m}x{n} k={k}: sigma rel err vs FP64 -- split-bf16 {err_mp:.2e}, "
          f"FP32 reference {err_f:.2e}; residual rel diff -- split-bf16 "
          f"{abs(rmp - r64) / r64:.2e}, FP32 reference {abs(rf - r64) / r64:.2e}")
    assert err_mp <= max(err_f, 1e-7)
    assert abs(rmp - r64) <= max(abs(rf - r64), 1e-7 * r64)


def test_blocked_qb_mixed_vs_fp32_reference(gpu_lib, oracle):
    rp = gpu_lib
    lib = rp.load()
    m, n, kstep, q = 6000, 1500, 64, 1
    j = np.arange(800, dtype=np.float64)
    sigma = np.where(j < 50, 1.0, np.exp(-(j - 50) / 40.0))
    sigma = sigma[sigma >= 1e-8]
    A64, A32 = _test_matrix(rp, m, n, sigma, 0.0, 1234)
    tol = 1e-6 * float(np.sqrt(np.sum(sigma**2)))
    seed = 55
    nmax = int(lib.rsvd_qb_max_cols(m, n, kstep, 0, 1))
    rn = rp.gaussian(seed, 0, n * nmax).astype(np.float32).astype(np.float64).reshape(nmax, n).T
    dRN64 = rp.DeviceMatrix.from_host(rn)
    dRN32 = rp.DeviceMatrix.from_host(rn.astype(np.float32))

    # FP64 QB (reference precision)
    Q64 = rp.DeviceMatrix(m, nmax)
    B64 = rp.DeviceMatrix(nmax, n)
    f64, nrm = C.c_int(0), C.c_double(0)
    assert lib.rsvd_dev_randqb(A64.ptr, m, n, m, kstep, 0, tol, q, 1, 1, seed, dRN64.ptr,
                               Q64.ptr, m, B64.ptr, nmax, C.byref(f64), C.byref(nrm)) == 0
    # mixed QB (B returned transposed)
    Qm = rp.DeviceMatrix(m, nmax, np.float32)
    Btm = rp.DeviceMatrix(n, nmax, np.float32)
    fmp = C.c_int(0)
    assert lib.rsvd_dev_randqb_mp(A32.ptr, m, n, m, kstep, 0, tol, q, 1, 1, seed, dRN32.ptr, 2,
                                  Qm.ptr, m, Btm.ptr, n, C.byref(fmp), C.byref(nrm)) == 0, \
        rp.last_error()
    Ah = A32.to_host().astype(np.float64)
    Qf, Bf, ff, _ = oracle.randQB_pb2(Ah.astype(np.float32), kstep, 0, tol, q, 1, omega=rn)

    def report(Q, B):
        Q, B = Q.astype(np.float64), B.astype(np.float64)
        res = np.linalg.norm(Ah - Q @ B)
        orth = np.abs(Q.T @ Q - np.eye(Q.shape[1])).max()
        return res, orth, np.linalg.svd(B, compute_uv=False) if orth < 1e-2 else None

    Q64h, B64h = Q64.to_host()[:, :f64.value], B64.to_host()[:f64.value]
    r64, o64, s64 = report(Q64h, B64h)
    rmp, omp_, smp = report(Qm.to_host()[:, :fmp.value], Btm.to_host()[:, :fmp.value].T)
    rf, of, _ = report(Qf, Bf)
    Uf, Sf, Vf = oracle.svd_fromQB_f32(Qf, Bf)
    sf = np.diag(Sf).astype(np.float64)
    kk = min(len(s64), len(smp), len(sf))
    dmp = np.abs(smp[:kk] - s64[:kk]).max()
    df = np.abs(sf[:kk] - s64[:kk]).max()
    print(f"blocked QB {m}x{n} TOL={tol:.3e}: rank FP64 {f64.value}, mixed {fmp.value}, "
          f"FP32 reference {ff}; ||A - QB|| FP64 {r64:.3e}, mixed {rmp:.3e}, FP32 reference "
          f"{rf:.3e}; max|Q^T Q - I| FP64 {o64:.1e}, mixed {omp_:.1e}, FP32 reference {of:.1e}; "
          f"max |d sigma| vs FP64 mixed {dmp:.2e}, FP32 reference {df:.2e}")
    assert rmp <= 1.05 * tol
    assert omp_ <= 1e-6 and omp_ <= of
    assert dmp <= max(df, 1e-7)
