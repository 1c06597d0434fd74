"""Learned (q, rank) front end: data, labels, training and evaluation.

The network and every kernel it needs live in librsvdpack (csrc/learned.hip);
this module drives them through the C ABI:

* family(i): matrix i of the power-law family -- m x n float (default 1024 x
  512), sigma_j = (j + 1)^-alpha with alpha ~ U[0.5, 3], plus Gaussian noise
  whose singular-value floor is nu = 10^U[-6, -2] (entry std nu / (sqrt(m) +
  sqrt(n))), Haar factors; alpha, nu and the seed follow from (seed, i).
* label(A): runs the plain adaptive method (blockrand in TOL mode, rra.c:239-381)
  for q = 0..4 and keeps the cheapest by a pass-count cost model,
  (2 q + 2) (frank_q + kstep): q* and rank* / min(m, n).
* train(...): Adam on the GPU with the library's forward / backward kernels.
* evaluate(...): learned path (one fixed-rank rSVD with the predicted q and
  rank -- a predicted rank at the sketch limit is solved as full rank, an
  exact thin QR of A and the small eigenproblem --, FP64 residual check,
  fallback) against the plain adaptive path at the same TOL.
"""
from __future__ import annotations

import ctypes as C
import sys
import time
from dataclasses import dataclass, field

import numpy as np

from . import rsvdpack as rp

M_DEF, N_DEF = 1024, 512
TOL_DEF = 1e-4
KSTEP_DEF = 32
P_DEF = 10
Q_PLAIN = 2


@dataclass
class FamilyMatrix:
    A64: rp.DeviceMatrix
    A32: rp.DeviceMatrix
    alpha: float
    nu: float
    seed: int


def family(i: int, seed: int, m: int = M_DEF, n: int = N_DEF) -> FamilyMatrix:
    rng = np.random.default_rng([seed, i])
    alpha = float(rng.uniform(0.5, 3.0))
    nu = float(10.0 ** rng.uniform(-6, -2))
    r = min(m, n)
    sigma = (1.0 + np.arange(r)) ** -alpha
    A64 = rp.DeviceMatrix(m, n)
    A32 = rp.DeviceMatrix(m, n, np.float32)
    lib = rp.load()
    lib.rsvd_dev_lowrank_test_matrix.restype = C.c_int
    lib.rsvd_dev_lowrank_test_matrix.argtypes = [C.c_longlong, C.c_longlong, C.c_int, C.c_void_p,
                                                 C.c_double, C.c_ulonglong, C.c_void_p,
                                                 C.c_longlong, C.c_void_p, C.c_longlong]
    rp._check(lib.rsvd_dev_lowrank_test_matrix(m, n, r, sigma.ctypes.data,
                                               nu / (np.sqrt(m) + np.sqrt(n)),
                                               (seed << 20) + i, A64.ptr, m, A32.ptr, m),
              "family matrix")
    return FamilyMatrix(A64, A32, alpha, nu, (seed << 20) + i)


def frob(A: rp.DeviceMatrix) -> float:
    """||A||_F of a device matrix (FP64, on the GPU)."""
    return float(np.linalg.norm(A.to_host()))


def plain_adaptive(A64: rp.DeviceMatrix, tol_rel: float, q: int, kstep: int = KSTEP_DEF,
                   p: int = P_DEF, seed: int = 1, anorm: float | None = None):
    """blockrand in TOL mode (k < 0): returns (frank, sigma, seconds)."""
    lib = rp.load()
    m, n = A64.rows, A64.cols
    if anorm is None:
        anorm = frob(A64)
    fr = C.c_int(0)
    u, s, v = C.c_void_p(), C.c_void_p(), C.c_void_p()
    lib.rsvd_sync()
    t0 = time.perf_counter()
    rp._check(lib.rsvd_dev_blockrand(A64.ptr, m, n, m, -1, p, tol_rel * anorm, 2, kstep, q, 1,
                                     seed, None, C.byref(fr), C.byref(u), C.byref(s),
                                     C.byref(v)), "rsvd_dev_blockrand")
    dt = time.perf_counter() - t0
    sig = np.empty(fr.value)
    if fr.value:
        rp._check(lib.rsvd_memcpy_d2h(sig.ctypes.data, s.value, 8 * fr.value), "d2h")
    for ptr in (u, s, v):
        lib.rsvd_dev_free(ptr.value)
    return fr.value, sig, dt


def label(fm: FamilyMatrix, tol_rel: float = TOL_DEF, kstep: int = KSTEP_DEF, p: int = P_DEF):
    """(q*, rank target, frank per q) from plain adaptive runs for q = 0..4.

    The rank target is log(frank) / log(min(m, n)) of the cheapest q.  Runs
    that miss TOL (TOL out of reach below full rank: the noise floor's energy
    alone exceeds it) cost as if they had gone to full rank; when every q
    misses, the target is 1 (full rank: the served path then solves such a
    matrix exactly)."""
    m, n = fm.A64.rows, fm.A64.cols
    mn = min(m, n)
    anorm = frob(fm.A64)
    franks, ok = [], []
    for q in range(5):
        fr, sig, _ = plain_adaptive(fm.A64, tol_rel, q, kstep, p, seed=fm.seed, anorm=anorm)
        res = np.sqrt(max(anorm ** 2 - float(np.sum(sig ** 2)), 0.0)) / anorm
        franks.append(fr)
        ok.append(res <= tol_rel * (1 + 1e-9))
    cost = [(2 * q + 2) * ((franks[q] if ok[q] else mn) + kstep) for q in range(5)]
    qs = int(np.argmin(cost))
    # the rank target on a log scale, log(frank) / log(min(m, n)): 1 = full rank
    return qs, (np.log(max(franks[qs], 1)) / np.log(mn) if ok[qs] else 1.0), franks


def features(fm: FamilyMatrix) -> np.ndarray:
    lib = rp.load()
    f = rp.DeviceMatrix(64, 1, np.float32)
    rp._check(lib.rsvd_dev_sketch_features(fm.A64.ptr, fm.A32.ptr, fm.A32.rows, fm.A32.cols,
                                           fm.A32.rows, fm.seed ^ 0x5EED5EED, f.ptr), "features")
    return f.to_host()[:, 0]


@dataclass
class Dataset:
    X: np.ndarray
    q: np.ndarray
    r: np.ndarray
    franks: list = field(default_factory=list)


def build_dataset(count: int, seed: int, tol_rel: float = TOL_DEF,
                  progress: int = 0) -> Dataset:
    """Features and labels of family matrices 0 .. count - 1; with progress > 0
    a line on stderr every `progress` matrices."""
    X, qs, rs, fr = [], [], [], []
    for i in range(count):
        if progress and i and i % progress == 0:
            print(f"[learned] dataset: {i} of {count} matrices", file=sys.stderr, flush=True)
        fm = family(i, seed)
        X.append(features(fm))
        q, r, franks = label(fm, tol_rel)
        qs.append(q)
        rs.append(r)
        fr.append(franks)
        del fm
    return Dataset(np.array(X, np.float32), np.array(qs, np.int32), np.array(rs, np.float32), fr)


def train(ds: Dataset, epochs: int = 400, batch: int = 64, lr: float = 2e-3, lam: float = 20.0,
          seed: int = 7):
    """Trains on the GPU; returns (theta device buffer, per-epoch mean loss)."""
    lib = rp.load()
    npar = lib.rsvd_mlp_num_params()
    theta = rp.DeviceMatrix(npar, 1, np.float32)
    rp._check(lib.rsvd_dev_mlp_init(seed, theta.ptr), "mlp init")
    dX = rp.DeviceMatrix.from_host(np.ascontiguousarray(ds.X).reshape(-1, 1))
    dq = rp.DeviceMatrix.from_host(np.ascontiguousarray(ds.q, np.int32).view(np.float32).reshape(-1, 1))
    dr = rp.DeviceMatrix.from_host(np.ascontiguousarray(ds.r).reshape(-1, 1))
    hist = np.zeros(epochs, np.float32)
    rp._check(lib.rsvd_dev_mlp_train(theta.ptr, dX.ptr, dq.ptr, dr.ptr, len(ds.q), lam, epochs,
                                     batch, lr, seed, hist.ctypes.data), "mlp train")
    return theta, hist


def learned_lowrank(theta: rp.DeviceMatrix, fm: FamilyMatrix, tol_rel: float = TOL_DEF,
                    p: int = P_DEF, margin: float = 0.1, kstep: int = KSTEP_DEF,
                    q_fallback: int = Q_PLAIN, keep: bool = False):
    """One learned solve: dict with frank, sigma, q_pred, k_pred, fell_back,
    resid_rel, seconds (and host U, V when keep)."""
    lib = rp.load()
    m, n = fm.A64.rows, fm.A64.cols
    fr, qp, kp, fb = C.c_int(0), C.c_int(0), C.c_int(0), C.c_int(0)
    res = C.c_double(0)
    u, s, v = C.c_void_p(), C.c_void_p(), C.c_void_p()
    lib.rsvd_sync()
    t0 = time.perf_counter()
    rp._check(lib.rsvd_dev_learned_lowrank(theta.ptr, fm.A64.ptr, fm.A32.ptr, m, n, m, tol_rel, p,
                                           margin, kstep, q_fallback, fm.seed, C.byref(fr),
                                           C.byref(u), C.byref(s), C.byref(v), C.byref(qp),
                                           C.byref(kp), C.byref(fb), C.byref(res)),
              "rsvd_dev_learned_lowrank")
    dt = time.perf_counter() - t0
    k = fr.value
    sig = np.empty(k)
    out = {"frank": k, "sigma": sig, "q_pred": qp.value, "k_pred": kp.value,
           "fell_back": int(fb.value), "resid_rel": res.value, "seconds": dt}
    if k:
        rp._check(lib.rsvd_memcpy_d2h(sig.ctypes.data, s.value, 8 * k), "d2h")
        if keep:
            U = np.empty((k, m))
            V = np.empty((k, n))
            rp._check(lib.rsvd_memcpy_d2h(U.ctypes.data, u.value, 8 * m * k), "d2h")
            rp._check(lib.rsvd_memcpy_d2h(V.ctypes.data, v.value, 8 * n * k), "d2h")
            out["U"], out["V"] = U.T, V.T
    for ptr in (u, s, v):
        lib.rsvd_dev_free(ptr.value)
    return out


def evaluate(theta, first: int, count: int, seed: int, tol_rel: float = TOL_DEF,
             margin: float = 0.1):
    """Learned vs plain adaptive (q = 2) on matrices first .. first + count - 1."""
    out = []
    for i in range(first, first + count):
        fm = family(i, seed)
        anorm = frob(fm.A64)
        fr_p, sig_p, t_p = plain_adaptive(fm.A64, tol_rel, Q_PLAIN, seed=fm.seed, anorm=anorm)
        res_p = np.sqrt(max(anorm ** 2 - float(np.sum(sig_p ** 2)), 0.0)) / anorm
        lr = learned_lowrank(theta, fm, tol_rel, margin=margin)
        out.append({"i": i, "plain_seconds": t_p, "plain_frank": fr_p, "plain_resid": res_p,
                    **{k: v for k, v in lr.items() if k != "sigma"}})
    return out
