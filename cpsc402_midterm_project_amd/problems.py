"""Synthetic systems of the benchmark configurations.

randsvd(n, kappa, seed, nrhs): A = U diag(sigma) V^T with U, V the Q factors of seeded Gaussian n x n matrices
(Householder QR, no sign normalisation, as the reference's CPU generator e1_scalar/three_precision_lab/
tpir/problems.py:8-37 does) and sigma_i = kappa^(-i/(n-1)); x_true ~ U(-1, 1) (n x nrhs); b = A x_true, all FP64.
Draw order from numpy.random.default_rng(seed): G_U, then G_V (standard normal, row-major n x n), then x_true.
For n >= 4096 the QR factorizations and products run on the GPU through torch when it is available (setup
only; the solver never uses torch); the recipe and the SHA-256 of the column-major bytes identify the matrix.
"""
from __future__ import annotations

import hashlib

import numpy as np


def _qr_q(G: np.ndarray, use_gpu: bool):
    if use_gpu:
        import torch
        t = torch.from_numpy(G).cuda()
        q, _ = torch.linalg.qr(t)
        return q
    return np.linalg.qr(G)[0]


def randsvd(n: int, kappa: float, seed: int, nrhs: int = 1, gpu: bool | None = None):
    """Returns (A, b, x_true, recipe) with A Fortran-ordered float64."""
    rng = np.random.default_rng(seed)
    if gpu is None:
        gpu = False
        if n >= 4096:
            try:
                import torch
                gpu = torch.cuda.is_available()
            except Exception:
                gpu = False
    G1 = rng.standard_normal((n, n))
    G2 = rng.standard_normal((n, n))
    x_true = rng.uniform(-1.0, 1.0, (n, nrhs))
    sigma = np.ones(n) if n == 1 else kappa ** (-np.arange(n, dtype=np.float64) / (n - 1))
    if gpu:
        import torch
        U = _qr_q(G1, True)
        del G1
        V = _qr_q(G2, True)
        del G2
        s = torch.from_numpy(sigma).cuda()
        A_t = (U * s) @ V.T
        del U, V
        xt = torch.from_numpy(x_true).cuda()
        b_t = A_t @ xt
        # column-major host copy: A^T in row-major order has the bytes of A in column-major order
        A = np.asfortranarray(A_t.cpu().numpy())
        b = np.asfortranarray(b_t.cpu().numpy())
        del A_t, b_t
        torch.cuda.empty_cache()
    else:
        U = _qr_q(G1, False)
        V = _qr_q(G2, False)
        A = np.asfortranarray((U * sigma) @ V.T)
        b = np.asfortranarray(A @ x_true)
    recipe = dict(family="randsvd", n=n, kappa=float(kappa), seed=int(seed), nrhs=int(nrhs),
                  generator="numpy.default_rng(seed): G_U, G_V standard_normal (n,n); x_true uniform(-1,1,(n,nrhs)); "
                            "Q of Householder QR; sigma_i = kappa^(-i/(n-1)); A = (U*sigma) V^T; b = A x_true",
                  qr_backend="torch-gpu" if gpu else "numpy-lapack")
    return A, b, np.asfortranarray(x_true), recipe


def sha256_colmajor(a: np.ndarray) -> str:
    return hashlib.sha256(np.asfortranarray(a).tobytes(order="F")).hexdigest()


def exact_lu_system(n: int, seed: int, fmt: str = "fp16"):
    """A = P^T L U with dyadic entries chosen so every step of a blocked LU with any of the trailing formats is
    exact: L unit lower with off-diagonal entries in {0, +-1/4, +-1/2} (so partial pivoting picks the diagonal
    with a strict margin), U upper with diagonal in {+-1, +-2} and off-diagonal small integers representable in
    the format.  Returns (A as float32 Fortran array, L, U, perm) with A[perm] = L @ U."""
    rng = np.random.default_rng(seed)
    L = np.tril(rng.choice([0.0, 0.25, -0.25, 0.5, -0.5], size=(n, n)), -1)
    # keep the Schur complements small: only a few nonzeros per column below the diagonal
    mask = rng.random((n, n)) < min(1.0, 4.0 / max(n, 1))
    L = L * np.tril(mask, -1) + np.eye(n)
    offd = [0.0, 1.0, -1.0, 2.0, -2.0] if fmt != "fp8" else [0.0, 1.0, -1.0, 0.5, -0.5]
    U = np.triu(rng.choice(offd, size=(n, n)), 1) * np.triu(rng.random((n, n)) < min(1.0, 4.0 / max(n, 1)), 1)
    U += np.diag(rng.choice([1.0, -1.0, 2.0, -2.0], size=n))
    perm = rng.permutation(n)
    PA = L @ U
    A = np.empty_like(PA)
    A[perm] = PA          # row i of P A is row perm[i] of A
    return np.asfortranarray(A.astype(np.float32)), L, U, perm
