"""The two large mixed-precision workloads at their full sizes.

* Fixed rank, 65536 x 16384 float32 A = U diag(sigma) V^T + N(0, 1e-4^2)
  noise, rank-2048 Haar factors, sigma_j = 1 / (j + 1); k = 512, p = 32, q = 2,
  s = 1, vnum = 2; split-bf16 passes (rsvd_dev_fixed_rank_mp precision 2) and
  fp16 passes (precision 6) against the FP64 plain path (rsvd_dev_fixed_rank) on the same A and Omega
  (reference algorithm: MKL/rank_revealing_algorithms_intel_mkl.c:73-234).
  Checks: every returned sigma within 1e-5 relative of the FP64 path's; the
  explicit FP64 residual ||A - U S V^T||_F within 1e-5 relative of the FP64
  path's; U^T U and V^T V within 5e-5 of I; the residual above the
  Eckart-Young bound sqrt(||A||^2 - sum_{j<k} (sigma_j + e)^2), where
  sigma_j + e bounds the j-th singular value of A from above (Weyl, e an upper
  bound of ||noise||_2: 1.1 * 1e-4 * (sqrt(m) + sqrt(n))).
* Blocked TOL mode, 100000 x 20000 float32 A = U diag(sigma) V^T, sigma = 1
  on a 300-value plateau, then exp(-(j - 300) / 150) down to 1e-8; kstep = 256,
  q = 1, s = 1, TOL = 1e-6 ||A||_F (low_rank_svd_blockrand_decomp_fixed_rank_or_prec,
  rank_revealing_algorithms_intel_mkl.c:239-381, QB :1576-1801); the mixed
  path (rsvd_dev_blockrand_mp, precision 2) against the FP64 path
  (rsvd_dev_blockrand).  Checks: the same reported rank; max |d sigma| <= 1e-6
  sigma_1 over the common values; the explicit FP64 residual of both between
  the Eckart-Young tail of the known spectrum and sqrt(tail^2 + TOL^2) (the
  QB leaves at most TOL outside range(Q)), up to the float rounding of A; the
  mixed one within TOL (in quadrature) of the FP64 one.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _resid(lib, rp, A64, m, n, k, U, ldu, S, V, ldv, f32):
    out = C.c_double(0)
    rc = lib.rsvd_dev_lowrank_residual(A64.ptr, m, n, m, k, U, ldu, S, V, ldv, f32,
                                       C.byref(out))
    assert rc == 0, rp.last_error()
    return out.value


def _gram_dev(A_host64):
    G = A_host64.T @ A_host64
    return float(np.abs(G - np.eye(G.shape[0])).max())


@pytest.mark.parametrize("prec", [2, 6])
def test_fixed_rank_65536x16384_vs_fp64(gpu_lib, prec):
    """prec 2: split-bf16 passes; 6: fp16 passes (two products)."""
    rp = gpu_lib
    lib = rp.load()
    m, n, r, k, p, q = 65536, 16384, 2048, 512, 32, 2
    sigma = 1.0 / (1.0 + np.arange(r, dtype=np.float64))
    A64 = rp.DeviceMatrix(m, n, np.float64)
    A32 = rp.DeviceMatrix(m, n, np.float32)
    assert lib.rsvd_dev_lowrank_test_matrix(m, n, r, sigma.ctypes.data, 1e-4, 20260101, A64.ptr,
                                            m, A32.ptr, m) == 0, rp.last_error()
    seed = 777
    U64, S64, V64 = rp.DeviceMatrix(m, k), rp.DeviceMatrix(k, 1), rp.DeviceMatrix(n, k)
    assert lib.rsvd_dev_fixed_rank(A64.ptr, m, n, m, k, p, 2, q, 1, seed, None, U64.ptr, m,
                                   S64.ptr, V64.ptr, n) == 0, rp.last_error()
    Um, Sm, Vm = (rp.DeviceMatrix(m, k, np.float32), rp.DeviceMatrix(k, 1),
                  rp.DeviceMatrix(n, k, np.float32))
    assert lib.rsvd_dev_fixed_rank_mp(A32.ptr, m, n, m, k, p, 2, q, 1, seed, None, prec, Um.ptr,
                                      m, Sm.ptr, Vm.ptr, n) == 0, rp.last_error()
    s64, smp = S64.to_host()[:, 0], Sm.to_host()[:, 0]
    assert np.all(np.diff(s64) >= 0) and np.all(np.diff(smp) >= 0)  # vnum = 2: ascending
    rel = np.abs(smp - s64) / s64
    print(f"precision {prec}: sigma rel err max {rel.max():.2e}")
    assert rel.max() <= 1e-5, rel.max()

    res64 = _resid(lib, rp, A64, m, n, k, U64.ptr, m, S64.ptr, V64.ptr, n, 0)
    resmp = _resid(lib, rp, A64, m, n, k, Um.ptr, m, Sm.ptr, Vm.ptr, n, 1)
    assert abs(resmp - res64) <= 1e-5 * res64, (resmp, res64)

    # ||A||_F: the residual of the rank-1 "approximation" with sigma = 0
    zero = rp.DeviceMatrix.from_host(np.zeros((1, 1)))
    anorm = _resid(lib, rp, A64, m, n, 1, U64.ptr, m, zero.ptr, V64.ptr, n, 0)
    e = 1.1 * 1e-4 * (np.sqrt(m) + np.sqrt(n))
    ey = np.sqrt(max(anorm**2 - np.sum((sigma[:k] + e) ** 2), 0.0))
    assert ey > 0.5 * res64          # the bound is not vacuous here
    assert resmp >= ey and res64 >= ey, (resmp, res64, ey)
    # the residual of this projection structure is sqrt(||A||^2 - sum sigma^2)
    assert abs(resmp**2 - (anorm**2 - np.sum(smp**2))) <= 1e-4 * anorm**2

    assert _gram_dev(Vm.to_host().astype(np.float64)) <= 5e-5
    assert _gram_dev(Um.to_host().astype(np.float64)) <= 5e-5


def test_blockrand_tol_100000x20000_split_bf16_vs_fp64(gpu_lib):
    rp = gpu_lib
    lib = rp.load()
    m, n = 100000, 20000
    j = np.arange(3100, dtype=np.float64)
    sigma = np.where(j < 300, 1.0, np.exp(-(j - 300) / 150.0))
    sigma = sigma[sigma >= 1e-8]
    r = len(sigma)
    A64 = rp.DeviceMatrix(m, n, np.float64)
    A32 = rp.DeviceMatrix(m, n, np.float32)
    assert lib.rsvd_dev_lowrank_test_matrix(m, n, r, sigma.ctypes.data, 0.0, 31337, A64.ptr, m,
                                            A32.ptr, m) == 0, rp.last_error()
    anorm = float(np.sqrt(np.sum(sigma**2)))
    tol = 1e-6 * anorm
    kstep, q, p = 256, 1, 10

    def run(mp):
        fr = C.c_int(0)
        u, sv, v = C.c_void_p(), C.c_void_p(), C.c_void_p()
        if mp:
            rc = lib.rsvd_dev_blockrand_mp(A32.ptr, m, n, m, -1, p, tol, 2, kstep, q, 1, 99, None,
                                           2, C.byref(fr), C.byref(u), C.byref(sv), C.byref(v))
        else:
            rc = lib.rsvd_dev_blockrand(A64.ptr, m, n, m, -1, p, tol, 2, kstep, q, 1, 99, None,
                                        C.byref(fr), C.byref(u), C.byref(sv), C.byref(v))
        assert rc == 0, rp.last_error()
        f = fr.value
        s = np.empty(f)
        assert lib.rsvd_memcpy_d2h(s.ctypes.data, sv.value, 8 * f) == 0
        res = _resid(lib, rp, A64, m, n, f, u.value, m, sv.value, v.value, n, 1 if mp else 0)
        for ptr in (u, sv, v):
            lib.rsvd_dev_free(ptr.value)
        return f, np.sort(s)[::-1], res

    f64, s64, res64 = run(False)
    fmp, smp, resmp = run(True)
    assert fmp == f64, (fmp, f64)
    assert 1500 < f64 < 3000
    kk = min(len(s64), len(smp))
    assert np.abs(smp[:kk] - s64[:kk]).max() <= 1e-6 * sigma[0]
    # A is exactly U diag(sigma) V^T, so the best rank-f residual is the
    # energy of the dropped values (Eckart-Young); the result is the rank-f
    # truncation of Q B with ||A - Q B|| <= TOL and (I - Q Q^T) A orthogonal to
    # Q B, so its residual^2 is at most tail^2 + TOL^2
    # (A as stored is that product rounded to float: a perturbation of norm
    # at most 2^-24 ||A||_F, which moves any residual by no more than that)
    tail = float(np.sqrt(np.sum(sigma[f64:] ** 2)))
    e_round = 2.0**-24 * anorm
    for res, slack in ((res64, 1.0), (resmp, 1.05)):
        assert res >= tail - e_round, (res, tail)
        assert res <= np.sqrt(tail**2 + (slack * tol) ** 2) + e_round, (res, tail, tol)
    # the FP64 path's QB leaves less than TOL outside range(Q), the mixed
    # path's no more than TOL either: the mixed residual is within TOL (in
    # quadrature) of the FP64 one
    assert resmp <= np.sqrt(res64**2 + (1.05 * tol) ** 2), (resmp, res64, tol)
    print(f"blocked: frank {f64}, residual FP64 {res64:.4e}, mixed {resmp:.4e}, tail {tail:.4e}, "
          f"TOL {tol:.3e}, max |d sigma| {np.abs(smp[:kk] - s64[:kk]).max():.2e}")
