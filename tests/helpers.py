"""Test matrices and comparison helpers shared by the test modules."""
from __future__ import annotations

import numpy as np


def haar(rng: np.random.Generator, m: int, n: int) -> np.ndarray:
    """m x n with orthonormal columns, Q of QR(N(0,1)) with sign(diag R) fixed."""
    X = rng.standard_normal((m, n))
    Q, R = np.linalg.qr(X)
    return Q * np.sign(np.diag(R))


def lowrank_matrix(m: int, n: int, sigma: np.ndarray, seed: int = 1) -> np.ndarray:
    """A = U diag(sigma) V^T with Haar U (m x r), V (n x r), r = len(sigma)."""
    rng = np.random.default_rng(seed)
    r = len(sigma)
    U = haar(rng, m, r)
    V = haar(rng, n, r)
    return (U * sigma) @ V.T


def config1_matrix() -> tuple[np.ndarray, np.ndarray]:
    """2000 x 1000, sigma_j = exp(-j/20), Haar factors (seeded)."""
    sigma = np.exp(-np.arange(1000) / 20.0)
    return lowrank_matrix(2000, 1000, sigma, seed=20240601), sigma


def align_signs(X: np.ndarray, Ref: np.ndarray) -> np.ndarray:
    """Flip the columns of X to best match the columns of Ref."""
    s = np.sign(np.sum(X * Ref, axis=0))
    s[s == 0] = 1
    return X * s


def projector_svd(A: np.ndarray, K: np.ndarray, k: int):
    """Top-k SVD of P_K A, P_K the orthogonal projector onto range(K)."""
    Qk, _ = np.linalg.qr(K)
    B = Qk.T @ A
    Ub, s, Vt = np.linalg.svd(B, full_matrices=False)
    return (Qk @ Ub)[:, :k], s[:k], Vt[:k].T


def rel(a, b) -> float:
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))
