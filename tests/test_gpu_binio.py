"""Binary matrix files loaded straight into HBM (rsvd_dev_load_binary) and
written from it (rsvd_dev_write_binary), in the reference's three layouts:
format 0 (MKL/mvf.c:77-133: int32 header, double entries, row after row),
format 1 (multi_core_mkl_code_64bit/matrix_vector_functions_intel_mkl.c:78-101:
int64 header, doubles) and format 2 (oneapi_code/matrix_vector_functions_one_api.c:71-95:
int64 header, floats).  Checks: the device matrix equals the file's entries
exactly (float entries widen exactly, doubles round to float as numpy does),
also with a leading dimension above the row count and across several blocks;
the writer reproduces the file byte for byte; and a fixed-rank rSVD of a
loaded file matches the oracle on the same Omega to 1e-10.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _file(path, A, fmt):
    hdr = np.int32(A.shape) if fmt == 0 else np.int64(A.shape)
    payload = A.astype("<f4" if fmt == 2 else "<f8")
    raw = hdr.tobytes() + np.ascontiguousarray(payload).tobytes()
    path.write_bytes(raw)
    return raw


@pytest.mark.parametrize("fmt", [0, 1, 2])
@pytest.mark.parametrize("dtype", [0, 1])
def test_load_exact_and_write_back(gpu_lib, tmp_path, fmt, dtype):
    rp = gpu_lib
    lib = rp.load()
    rng = np.random.default_rng(10 * fmt + dtype)
    m, n = (5000, 1100) if fmt != 2 else (9000, 1100)   # > 32 MiB: two or more blocks
    A = rng.standard_normal((m, n))
    f = tmp_path / f"f{fmt}.bin"
    raw = _file(f, A, fmt)
    npt = np.float32 if dtype == 1 else np.float64
    want = (A.astype(np.float32) if fmt == 2 else A).astype(npt)
    ld = m + 3
    D = rp.DeviceMatrix(ld, n, npt)
    assert lib.rsvd_dev_load_binary(str(f).encode(), fmt, dtype, D.ptr, ld) == 0, rp.last_error()
    got = D.to_host()[:m]
    assert np.array_equal(got, want)
    g = tmp_path / f"g{fmt}.bin"
    assert lib.rsvd_dev_write_binary(str(g).encode(), fmt, dtype, D.ptr, m, n, ld) == 0, \
        rp.last_error()
    if dtype == 0 or fmt == 2:
        assert g.read_bytes() == raw
    else:  # double file through a float matrix: the header and float-rounded entries
        back = np.frombuffer(g.read_bytes()[16 if fmt else 8:], dtype="<f8").reshape(m, n)
        assert np.array_equal(back, A.astype(np.float32).astype(np.float64))


def test_load_odd_shapes_and_errors(gpu_lib, tmp_path):
    rp = gpu_lib
    lib = rp.load()
    for shape in [(1, 1), (100000, 1), (1, 70000), (37, 33)]:
        A = np.arange(shape[0] * shape[1], dtype=np.float64).reshape(shape) / 7.0
        f = tmp_path / "o.bin"
        _file(f, A, 1)
        D = rp.DeviceMatrix(shape[0], shape[1])
        assert lib.rsvd_dev_load_binary(str(f).encode(), 1, 0, D.ptr, shape[0]) == 0, \
            rp.last_error()
        assert np.array_equal(D.to_host(), A)
    f = tmp_path / "t.bin"
    raw = _file(f, np.ones((4000, 3000)), 0)
    f.write_bytes(raw[:-100])
    D = rp.DeviceMatrix(4000, 3000)
    assert lib.rsvd_dev_load_binary(str(f).encode(), 0, 0, D.ptr, 4000) != 0
    assert "truncated" in rp.last_error()
    assert lib.rsvd_dev_load_binary(str(f).encode(), 0, 0, D.ptr, 10) != 0  # lda < rows
    # the library still works after the failures
    f2 = tmp_path / "ok.bin"
    _file(f2, np.eye(5), 0)
    E = rp.DeviceMatrix(5, 5)
    assert lib.rsvd_dev_load_binary(str(f2).encode(), 0, 0, E.ptr, 5) == 0
    assert np.array_equal(E.to_host(), np.eye(5))


def test_rsvd_on_loaded_file_matches_oracle(gpu_lib, oracle, tmp_path):
    """A reference-style driver: the matrix comes from a format-0 file."""
    rp = gpu_lib
    lib = rp.load()
    from tests.helpers import lowrank_matrix
    m, n, k, p, q, seed = 2000, 1000, 100, 10, 2, 4040
    A = lowrank_matrix(m, n, np.exp(-np.arange(1000) / 20.0), seed=77)
    f = tmp_path / "drv.bin"
    _file(f, A, 0)
    mm, nn = C.c_longlong(0), C.c_longlong(0)
    assert lib.rsvd_binary_file_shape(str(f).encode(), 0, C.byref(mm), C.byref(nn)) == 0
    D = rp.DeviceMatrix(mm.value, nn.value)
    assert lib.rsvd_dev_load_binary(str(f).encode(), 0, 0, D.ptr, m) == 0, rp.last_error()
    U, s, V = rp.dev_fixed_rank(D, k, p, 2, q, 1, seed)
    Ur, Sr, Vr = oracle.fixed_rank(A, k, p, 2, q, 1, seed=seed)
    sr = np.diag(Sr)
    assert np.max(np.abs(s - sr) / sr) < 1e-10
    res = np.linalg.norm(A - (U * s) @ V.T)
    resr = np.linalg.norm(A - Ur @ Sr @ Vr.T)
    assert abs(res - resr) < 1e-10 * resr


def test_load_rate(gpu_lib, tmp_path):
    """A 1.6 GB format-0 file into HBM: report the rate (file in page cache)."""
    import time
    rp = gpu_lib
    lib = rp.load()
    m, n = 40000, 5000
    rng = np.random.default_rng(1)
    A = rng.standard_normal((m, n))
    f = tmp_path / "big.bin"
    _file(f, A, 0)
    del A
    D = rp.DeviceMatrix(m, n)
    t0 = time.perf_counter()
    assert lib.rsvd_dev_load_binary(str(f).encode(), 0, 0, D.ptr, m) == 0, rp.last_error()
    dt = time.perf_counter() - t0
    print(f"loaded {8 * m * n / 1e9:.2f} GB in {dt:.3f} s: {8 * m * n / dt / 1e9:.2f} GB/s")
    with open(f, "rb") as fh:
        fh.seek(8)
        row = np.frombuffer(fh.read(8 * n), dtype="<f8")
    assert np.array_equal(D.to_host()[0], row)
