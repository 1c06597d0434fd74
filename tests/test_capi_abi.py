"""librsvdpack.so loads and exports every function the headers declare (CPU only).

No compute call is made here: without a GPU the library must refuse to run
(checked in a subprocess so the parent never creates a HIP context).
"""
import ctypes
import re
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
LIB = ROOT / "lowrankmatrixdecompositioncodes_amd" / "librsvdpack.so"


def declared_functions():
    names = []
    for h in sorted((ROOT / "include").glob("*.h")):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(([^;{}]*)\)\s*;", text):
            name = m.group(1)
            if name in {"if", "for", "while", "return", "sizeof", "min", "max"}:
                continue
            names.append(name)
    return sorted(set(names))


def test_headers_declare_reference_api():
    names = declared_functions()
    for must in ["low_rank_svd_rand_decomp_fixed_rank",
                 "low_rank_svd_blockrand_decomp_fixed_rank_or_prec",
                 "randQB_pb", "randQB_pb_new", "matrix_new", "matrix_delete",
                 "initialize_random_matrix", "QR_factorization_getQ",
                 "use_low_rank_svd_for_approximation", "rsvd_last_error"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    assert LIB.exists(), "build librsvdpack.so first (__graft_entry__.build())"
    out = subprocess.run(["nm", "-D", "--defined-only", str(LIB)], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing


def test_library_loads_and_host_storage_works():
    lib = ctypes.CDLL(str(LIB))
    lib.matrix_new.restype = ctypes.c_void_p
    lib.matrix_new.argtypes = [ctypes.c_int, ctypes.c_int]
    lib.matrix_delete.argtypes = [ctypes.c_void_p]
    lib.matrix_get_element.restype = ctypes.c_double
    lib.matrix_get_element.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    lib.matrix_set_element.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_double]
    M = lib.matrix_new(3, 4)
    assert lib.matrix_get_element(M, 2, 3) == 0.0  # calloc'ed, as mvf.c:12
    lib.matrix_set_element(M, 2, 1, 7.5)
    assert lib.matrix_get_element(M, 2, 1) == 7.5
    lib.matrix_delete(M)


def test_binary_file_roundtrip(tmp_path):
    """mvf.c:77-133 format: int32 m, int32 n, then m*n doubles row-major."""
    import numpy as np
    from lowrankmatrixdecompositioncodes_amd import rsvdpack as rp
    lib = rp.load()
    A = np.arange(12.0).reshape(3, 4) + 0.25
    f = tmp_path / "a.bin"
    raw = np.int32([3, 4]).tobytes() + A.astype("<f8").tobytes()  # row-major payload
    f.write_bytes(raw)
    M = lib.matrix_load_from_binary_file(str(f).encode())
    B = rp.from_mat(M, free=False)
    assert np.array_equal(A, B)
    g = tmp_path / "b.bin"
    lib.matrix_write_to_binary_file(M, str(g).encode())
    assert g.read_bytes() == raw
    lib.matrix_delete(M)


def test_resize_and_gather_semantics():
    import numpy as np
    from lowrankmatrixdecompositioncodes_amd import rsvdpack as rp
    lib = rp.load()
    A = np.arange(20.0).reshape(4, 5)
    for fn, want in [("resize_matrix_by_columns", A[:, :2]),
                     ("resize_matrix_by_columns_from_end", A[:, -2:]),
                     ("resize_matrix_by_rows", A[:2]),
                     ("resize_matrix_by_rows_from_end", A[-2:])]:
        M = rp.to_mat(A)
        getattr(lib, fn)(ctypes.byref(M), 2)
        assert np.array_equal(rp.from_mat(M), want), fn
    M = rp.to_mat(A)
    idx = (ctypes.c_int * 2)(4, 1)
    Mc = lib.matrix_new(4, 2)
    lib.matrix_get_selected_columns(M, idx, Mc)
    assert np.array_equal(rp.from_mat(Mc), A[:, [4, 1]])
    Mr = lib.matrix_new(2, 5)
    lib.matrix_get_selected_rows(M, (ctypes.c_int * 2)(3, 0), Mr)
    assert np.array_equal(rp.from_mat(Mr), A[[3, 0]])
    lib.matrix_delete(M)


def test_compute_without_gpu_fails_loudly():
    """No CPU fallback: with no visible GPU the compute entry points refuse."""
    code = (
        "import numpy as np\n"
        "from lowrankmatrixdecompositioncodes_amd import rsvdpack as rp\n"
        "try:\n"
        "    rp.low_rank_svd_rand_decomp_fixed_rank(np.eye(20), 3, 2, 2, 2, 1)\n"
        "    print('RAN')\n"
        "except rp.RsvdError as e:\n"
        "    print('ERR', e)\n"
    )
    env = {"HIP_VISIBLE_DEVICES": "-1", "ROCR_VISIBLE_DEVICES": "-1", "PATH": "/usr/bin:/bin",
           "PYTHONPATH": str(ROOT)}
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env,
                       cwd=str(ROOT), timeout=120)
    assert "ERR" in r.stdout, r.stdout + r.stderr
