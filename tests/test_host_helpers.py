"""The L1 host helpers of the C API against their definitions
(multi_core_mkl_code/matrix_vector_functions_intel_mkl.c), and the two QB / SVD
approximation reporters (rank_revealing_algorithms_intel_mkl.c:2337-2347,
2388-2398).

The helpers are plain host loops in the reference; each test checks the
element rule the reference states, on matrices whose every entry is distinct
so a transposed index or an off-by-one fails:
* matrix_copy (mvf.c:216-224): D = S element for element, D's extra storage
  untouched;
* matrix_build_transpose (:246-253): Mt(j, i) = M(i, j), non-square;
* matrix_sub (:274-283): A = A - B in place, B unchanged;
* initialize_diagonal_matrix (:490-499): D(i, i) = data(i), off-diagonal kept;
* invert_diagonal_matrix (:519-528): Dinv(i, i) = 1 / D(i, i), only the
  diagonal written (the reference never clears Dinv);
* matrix_copy_symmetric (:757-768): S(i, j) = M(i, j) for j >= i, the strict
  lower triangle of S untouched, rectangular M included.
"""

import numpy as np
import pytest

from lowrankmatrixdecompositioncodes_amd import rsvdpack as rp


@pytest.fixture(scope="module")
def lib():
    return rp.load()


def _distinct(m, n, seed):
    rng = np.random.default_rng(seed)
    return rng.permutation(m * n).reshape(m, n).astype(np.float64) + 0.5


def test_matrix_copy(lib):
    S = _distinct(5, 3, 1)
    Sm, Dm = rp.to_mat(S), rp.to_mat(np.full((5, 3), -7.0))
    lib.matrix_copy(Dm, Sm)
    assert np.array_equal(rp.from_mat(Dm), S)
    assert np.array_equal(rp.from_mat(Sm), S)


def test_matrix_build_transpose(lib):
    M = _distinct(4, 7, 2)
    Mm, Tm = rp.to_mat(M), lib.matrix_new(7, 4)
    lib.matrix_build_transpose(Tm, Mm)
    assert np.array_equal(rp.from_mat(Tm), M.T)
    lib.matrix_delete(Mm)


def test_matrix_sub(lib):
    A, B = _distinct(6, 4, 3), _distinct(6, 4, 4) * 0.25
    Am, Bm = rp.to_mat(A), rp.to_mat(B)
    lib.matrix_sub(Am, Bm)
    assert np.array_equal(rp.from_mat(Am), A - B)
    assert np.array_equal(rp.from_mat(Bm), B)


def test_initialize_and_invert_diagonal(lib):
    n = 6
    data = np.array([2.0, -4.0, 0.5, 8.0, -0.25, 3.0])
    v = lib.vector_new(n)
    for i, x in enumerate(data):
        lib.vector_set_element(v, i, x)
    base = _distinct(n, n, 5)
    Dm = rp.to_mat(base)
    lib.initialize_diagonal_matrix(Dm, v)
    D = rp.from_mat(Dm, free=False)
    want = base.copy()
    np.fill_diagonal(want, data)
    assert np.array_equal(D, want)
    # only the diagonal of Dinv is written
    junk = _distinct(n, n, 6)
    Im = rp.to_mat(junk)
    lib.invert_diagonal_matrix(Im, Dm)
    want = junk.copy()
    np.fill_diagonal(want, 1.0 / data)
    assert np.array_equal(rp.from_mat(Im), want)
    lib.matrix_delete(Dm)
    lib.vector_delete(v)


@pytest.mark.parametrize("shape", [(5, 5), (3, 6), (6, 3)])
def test_matrix_copy_symmetric(lib, shape):
    m, n = shape
    M = _distinct(m, n, 7)
    base = -_distinct(m, n, 8)
    Mm, Sm = rp.to_mat(M), rp.to_mat(base)
    lib.matrix_copy_symmetric(Sm, Mm)
    i, j = np.indices((m, n))
    want = np.where(j >= i, M, base)
    assert np.array_equal(rp.from_mat(Sm), want)
    lib.matrix_delete(Mm)


def _report(out):
    """The three numbers of one report (norm(M), norm(P), percent error)."""
    vals = {}
    for line in out.splitlines():
        for key in ("norm(M,fro) = ", "norm(P,fro) = ", "percent error = "):
            if line.startswith(key):
                vals[key] = float(line[len(key):])
    return [vals[k] for k in ("norm(M,fro) = ", "norm(P,fro) = ", "percent error = ")]


@pytest.mark.gpu
def test_qb_and_svd_reports(gpu_lib, capfd):
    """use_QB_decomp_for_approximation and use_low_rank_svd_for_approximation
    print norm(M), norm(P) and 100 norm(M - P) / norm(M) for P = Q B and
    P = U S V^T, each with "%f" (six decimals)."""
    lib = gpu_lib.load()
    rng = np.random.default_rng(11)
    M = rng.standard_normal((90, 60))
    Q, _ = np.linalg.qr(rng.standard_normal((90, 12)))
    B = Q.T @ M
    Mm, Qm, Bm = (gpu_lib.to_mat(x) for x in (M, Q, B))
    capfd.readouterr()
    lib.use_QB_decomp_for_approximation(Mm, Qm, Bm)
    got = _report(capfd.readouterr().out)
    P = Q @ B
    want = [np.linalg.norm(M), np.linalg.norm(P),
            100 * np.linalg.norm(M - P) / np.linalg.norm(M)]
    assert np.allclose(got, want, rtol=0, atol=1.5e-6), (got, want)
    U, s, Vt = np.linalg.svd(M, full_matrices=False)
    k = 9
    Um, Sm, Vm = (gpu_lib.to_mat(x) for x in (U[:, :k], np.diag(s[:k]), Vt[:k].T))
    lib.use_low_rank_svd_for_approximation(Mm, Um, Sm, Vm)
    got = _report(capfd.readouterr().out)
    want = [np.linalg.norm(M), np.sqrt(np.sum(s[:k] ** 2)),
            100 * np.sqrt(np.sum(s[k:] ** 2)) / np.linalg.norm(M)]
    assert np.allclose(got, want, rtol=0, atol=1.5e-6), (got, want)
    for x in (Mm, Qm, Bm, Um, Sm, Vm):
        lib.matrix_delete(x)
