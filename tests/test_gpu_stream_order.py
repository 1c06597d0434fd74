"""Stream-order contract of the device-pointer entry points.

The library runs on its own non-blocking HIP stream.  Every rsvd_dev_* routine
(1) orders its work after what the caller queued on the null stream before the
call and (2) returns only when its outputs are complete.  Both are checked with
HIP's own runtime through ctypes, outside the library's stream:

* outputs read with a plain hipMemcpy right after the call equal the outputs
  read through the library's stream after rsvd_sync();
* an input overwritten by an asynchronous device-to-device copy on the null
  stream just before the call is the input the call sees.
"""
import ctypes as C

import numpy as np
import pytest

from tests.helpers import lowrank_matrix

pytestmark = pytest.mark.gpu

D2H, D2D = 2, 3


@pytest.fixture(scope="module")
def hip():
    lib = C.CDLL("libamdhip64.so")
    lib.hipMemcpy.restype = C.c_int
    lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    lib.hipMemcpyAsync.restype = C.c_int
    lib.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
    lib.hipDeviceSynchronize.restype = C.c_int
    return lib


def _hip_read(hip, d, shape, dtype):
    out = np.empty(shape, dtype=dtype, order="F")
    assert hip.hipMemcpy(out.ctypes.data, d.ptr, out.nbytes, D2H) == 0
    return out


def test_outputs_complete_on_return(gpu_lib, hip):
    rp = gpu_lib
    lib = rp.load()
    m, n, k, p = 3000, 1200, 64, 16
    sig = 1.0 / (1.0 + np.arange(400))
    A = lowrank_matrix(m, n, sig, seed=11).astype(np.float32)
    dA = rp.DeviceMatrix.from_host(A)
    U, S, V = (rp.DeviceMatrix(m, k, np.float32), rp.DeviceMatrix(k, 1),
               rp.DeviceMatrix(n, k, np.float32))
    rc = lib.rsvd_dev_fixed_rank_mp(dA.ptr, m, n, m, k, p, 2, 2, 1, 5, None, 2, U.ptr, m, S.ptr,
                                    V.ptr, n)
    assert rc == 0, rp.last_error()
    # read at once with the runtime's blocking copy (null stream)
    s_now = _hip_read(hip, S, (k, 1), np.float64)
    v_now = _hip_read(hip, V, (n, k), np.float32)
    assert lib.rsvd_sync() == 0
    np.testing.assert_array_equal(s_now, S.to_host())
    np.testing.assert_array_equal(v_now, V.to_host())
    assert np.all(s_now > 0)


def test_input_written_on_null_stream_is_seen(gpu_lib, hip):
    rp = gpu_lib
    lib = rp.load()
    m, n, k, p = 1500, 700, 20, 8
    A1 = lowrank_matrix(m, n, np.exp(-np.arange(300) / 10.0), seed=3)
    A2 = lowrank_matrix(m, n, 2.0 * np.exp(-np.arange(300) / 30.0), seed=4)
    dA, dA2 = rp.DeviceMatrix.from_host(A1), rp.DeviceMatrix.from_host(A2)

    def run(d):
        U, S, V = rp.DeviceMatrix(m, k), rp.DeviceMatrix(k, 1), rp.DeviceMatrix(n, k)
        assert lib.rsvd_dev_fixed_rank(d.ptr, m, n, m, k, p, 2, 2, 1, 9, None, U.ptr, m, S.ptr,
                                       V.ptr, n) == 0, rp.last_error()
        return S.to_host()[:, 0]

    s1 = run(dA)
    s_expect = run(dA2)
    # queue A <- A2 on the null stream, then call at once
    assert hip.hipMemcpyAsync(dA.ptr, dA2.ptr, dA.nbytes, D2D, None) == 0
    s2 = run(dA)
    assert hip.hipDeviceSynchronize() == 0
    np.testing.assert_array_equal(s2, s_expect)
    assert np.abs(s1 - s2).max() > 0.1
