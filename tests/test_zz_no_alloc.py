"""rsvd_dev_fixed_rank_noalloc: the allocation-free variant of rsvd_dev_fixed_rank.

It takes every device buffer from the library's static scratch array.  Checked
through the C ABI on the GPU:
* two calls in a row, the second on a smaller matrix, each against
  rsvd_dev_fixed_rank on the same A and Omega (FP64; singular values and the
  residual within 1e-12 relative, U and V up to column signs within 1e-9);
* no device memory is taken from the runtime during those calls (the free
  device memory the runtime reports does not shrink, within 1 MiB, after the
  library's cache was released);
* a problem beyond the scratch array is refused with an error that says so,
  and the library keeps working afterwards.
"""
import ctypes as C

import numpy as np
import pytest

from tests.helpers import align_signs, lowrank_matrix

pytestmark = pytest.mark.gpu


def _outputs(rp, m, n, k):
    return rp.DeviceMatrix(m, k), rp.DeviceMatrix(k, 1), rp.DeviceMatrix(n, k)


def _call(rp, fn, dA, m, n, k, p, vnum, q, seed, out):
    U, S, V = out
    rc = fn(dA.ptr, m, n, m, k, p, vnum, q, 1, seed, None, U.ptr, m, S.ptr, V.ptr, n)
    assert rc == 0, rp.last_error()


def _host(out):
    U, S, V = out
    return U.to_host(), S.to_host()[:, 0], V.to_host()


def _run(rp, fn, dA, m, n, k, p, vnum, q, seed):
    out = _outputs(rp, m, n, k)
    _call(rp, fn, dA, m, n, k, p, vnum, q, seed, out)
    return _host(out)


def _free_mem(hip):
    free, total = C.c_size_t(0), C.c_size_t(0)
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


@pytest.fixture(scope="module")
def hip():
    lib = C.CDLL("libamdhip64.so")
    lib.hipMemGetInfo.restype = C.c_int
    lib.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    return lib


def test_noalloc_matches_fixed_rank_twice(gpu_lib, hip):
    rp = gpu_lib
    lib = rp.load()
    cases = [  # (m, n, k, p, vnum, q): the second call smaller than the first
        (3000, 1400, 96, 16, 2, 2),
        (900, 500, 30, 10, 1, 1),
    ]
    mats = []
    for i, (m, n, k, p, vnum, q) in enumerate(cases):
        A = lowrank_matrix(m, n, np.exp(-np.arange(min(m, n)) / 25.0), seed=40 + i)
        mats.append((A, rp.DeviceMatrix.from_host(A)))
    # the existing entry point first (its buffers come from the pool)
    want = [_run(rp, lib.rsvd_dev_fixed_rank, dA, *cs, 300 + i)
            for i, (cs, (_, dA)) in enumerate(zip(cases, mats))]
    outs = [_outputs(rp, cs[0], cs[1], cs[2]) for cs in cases]
    lib.rsvd_release_cache()
    assert lib.rsvd_sync() == 0
    free0 = _free_mem(hip)
    for i, (cs, (_, dA), out) in enumerate(zip(cases, mats, outs)):
        _call(rp, lib.rsvd_dev_fixed_rank_noalloc, dA, *cs, 300 + i, out)
    free1 = _free_mem(hip)
    # the library took no device memory from the runtime during the calls
    # (its panels alone, 8 * 112 * 3000 bytes for Y, would show here)
    assert free0 - free1 < (1 << 20), (free0, free1)
    got = [_host(out) for out in outs]
    for (m, n, k, p, vnum, q), (A, _), (Uw, sw, Vw), (Ug, sg, Vg) in zip(cases, mats, want, got):
        np.testing.assert_allclose(sg, sw, rtol=1e-12)
        np.testing.assert_allclose(align_signs(Ug, Uw), Uw, atol=1e-9)
        np.testing.assert_allclose(align_signs(Vg, Vw), Vw, atol=1e-9)
        rw = np.linalg.norm(A - (Uw * sw) @ Vw.T)
        rg = np.linalg.norm(A - (Ug * sg) @ Vg.T)
        assert abs(rg - rw) <= 1e-12 * np.linalg.norm(A)


def test_noalloc_refuses_beyond_scratch(gpu_lib):
    rp = gpu_lib
    lib = rp.load()
    # 8 (k + p) (3 n + 2 m) bytes of panels alone: 8 * 600 * (3 * 20000 + 2 * 40000) > 256 MiB
    m, n, k, p = 40000, 20000, 580, 20
    dA = rp.DeviceMatrix(1, 1)
    rc = lib.rsvd_dev_fixed_rank_noalloc(dA.ptr, m, n, m, k, p, 2, 1, 1, 1, None, dA.ptr, m,
                                         dA.ptr, dA.ptr, n)
    assert rc != 0
    assert "scratch" in rp.last_error()
    # the library still works
    A = lowrank_matrix(400, 300, np.exp(-np.arange(300) / 10.0), seed=8)
    dA2 = rp.DeviceMatrix.from_host(A)
    _, s, _ = _run(rp, lib.rsvd_dev_fixed_rank_noalloc, dA2, 400, 300, 20, 5, 2, 2, 12)
    _, s_ref, _ = _run(rp, lib.rsvd_dev_fixed_rank, dA2, 400, 300, 20, 5, 2, 2, 12)
    np.testing.assert_allclose(s, s_ref, rtol=1e-12)
