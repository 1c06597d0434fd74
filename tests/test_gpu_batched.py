"""Batched fixed-rank rSVD (rsvd_dev_fixed_rank_batched) against the oracle.

Each matrix of a small uneven batch (shapes drawn like the 4096-matrix batch
workload, smaller) is decomposed by the batched GPU path and, on the CPU, by
the oracle (FP64 restatement of low_rank_svd_rand_decomp_fixed_rank, vnum = 2)
with the same Omega (seeded stream, seed + i).  The batched passes over A
run in split bf16 (fp32 accuracy) and the small solves in FP64, so agreement
is at the fp32 level: singular values to 2e-5 of sigma_0 (= 1) absolute and
the residual ||A - U S V^T||_F to 1e-3 relative (of a residual that is itself
small, the rSVD's approximation error).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def make_batch(rp, count, mmax, nmax, seed):
    rng = np.random.default_rng(seed)
    m = np.exp(rng.uniform(np.log(64), np.log(mmax), count)).astype(np.int32)
    n = np.array([int(np.exp(rng.uniform(np.log(32), np.log(min(mi, nmax))))) for mi in m],
                 dtype=np.int32)
    k = np.minimum(32, n // 2).astype(np.int32)
    tau = rng.uniform(4, 64, count)
    b = rp.Batch(m, n, k, 8)
    b.fill_test_matrices(tau, seed)
    return b


@pytest.fixture(scope="module")
def batch(gpu_lib):
    return make_batch(gpu_lib, 40, 700, 400, 11)


def test_batched_matches_oracle(gpu_lib, oracle, batch):
    st = batch.run(q=2, seed=900)
    assert st == 0
    res = batch.results()
    worst_s, worst_r = 0.0, 0.0
    for i in range(batch.count):
        A = batch.matrix(i).astype(np.float64)
        m, n, k = A.shape[0], A.shape[1], int(batch.k[i])
        l = k + batch.p
        om = gpu_lib.gaussian(900 + i, 0, n * l).reshape(l, n).T
        Ur, Sr, Vr = oracle.fixed_rank(A, k, batch.p, 2, 2, 1, omega=om)
        sr = np.diag(Sr)
        u, s, v = res[i]
        assert np.all(np.diff(s) >= 0), "vnum = 2 order (ascending)"
        worst_s = max(worst_s, np.abs(s - sr).max())
        r_gpu = np.linalg.norm(A - (u.astype(np.float64) * s) @ v.astype(np.float64).T)
        r_ref = np.linalg.norm(A - Ur @ Sr @ Vr.T)
        worst_r = max(worst_r, abs(r_gpu - r_ref) / max(r_ref, 1e-30))
    print(f"batched vs oracle: max |sigma - sigma_ref| = {worst_s:.2e}, "
          f"max residual rel diff = {worst_r:.2e}; tuner: {gpu_lib.batched_last_config()}")
    assert worst_s < 2e-5
    assert worst_r < 1e-3


def test_batched_replay_is_deterministic(gpu_lib, batch):
    batch.run(q=2, seed=900)
    a = [x[1] for x in batch.results()]
    batch.run(q=2, seed=900)  # graph replay
    b = [x[1] for x in batch.results()]
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_batched_q1_and_orthonormal_factors(gpu_lib):
    b = make_batch(gpu_lib, 12, 300, 200, 5)
    assert b.run(q=1, seed=3) == 0
    for u, s, v in b.results():
        k = len(s)
        assert np.abs(u.T.astype(np.float64) @ u - np.eye(k)).max() < 1e-4
        assert np.all(s > 0)


@pytest.mark.parametrize("variant", ["1,1", "2,2", "4,4", "5,5"])
def test_batched_pass_variants_agree(gpu_lib, batch, monkeypatch, variant):
    """Every pass variant the autotuner chooses from (rows per wave 1 / 2 / 4,
    and 5 = four rows with the k range split over the workgroup's waves),
    fixed through RSVD_BATCHED_PASS, gives the tuned run's singular values to
    the fp32 level."""
    lib = gpu_lib.load()
    lib.rsvd_release_cache()
    assert batch.run(q=2, seed=900) == 0
    want = [x[1] for x in batch.results()]
    monkeypatch.setenv("RSVD_BATCHED_PASS", variant)
    lib.rsvd_release_cache()
    try:
        assert batch.run(q=2, seed=900) == 0
        cfg = gpu_lib.batched_last_config()
        got = [x[1] for x in batch.results()]
    finally:
        monkeypatch.delenv("RSVD_BATCHED_PASS")
        lib.rsvd_release_cache()
    nn, tn = variant.split(",")
    assert f"rt_nn={nn} rt_tn={tn}" in cfg, cfg
    worst = max(np.abs(g - w).max() for g, w in zip(got, want))
    assert worst < 2e-6, worst


def test_batched_rejects_bad_sizes(gpu_lib):
    b = gpu_lib.Batch([100, 100], [60, 60], [45, 10], 8)  # k + p = 53 > 48
    with pytest.raises(gpu_lib.RsvdError):
        b.run(q=2, seed=1)
