"""FP64 restatement of the reference's Ritz-vector training right-hand sides (test
infrastructure only).

lanczos          lib/create_dataset.py:28-52  (_lanczos_algorithm): Lanczos from the
                 normalised rhs with a three-term recurrence and, with ortho_iters = inf,
                 reorthogonalisation of w against v_{j-2} .. v_0 (newest first); stops
                 early when beta < cut_off_tol.
create_ritz_vec  lib/create_dataset.py:54-68: eigh_tridiagonal of (alpha, beta), Ritz
                 vectors (V^T Q[:, :num])^T, ascending Ritz values.
training_rhs     preprocess.py:38-58 (createTrainingData): coefficients N(0,1), the first
                 60 % (the smallest Ritz values) scaled by 9, b = coef^T R, each b
                 normalised.  The coefficient matrix is an argument here (the reference
                 draws it with np.random inside) so that a GPU run can use the same one.
"""
from __future__ import annotations

import numpy as np
import scipy.linalg


def lanczos(A, init_v, m, ortho_iters=np.inf, cut_off_tol=1e-10):
    """-> (V [m, n] rows orthonormal, alpha [m], beta [m-1]); truncated at a cut-off."""
    n = A.shape[0]
    V = np.zeros((m, n))
    alpha = np.zeros(m)
    beta = np.zeros(m - 1)
    V[0] = init_v / np.linalg.norm(init_v)
    w = A @ V[0]
    alpha[0] = w.dot(V[0])
    w = w - alpha[0] * V[0]
    for j in range(1, m):
        beta[j - 1] = np.linalg.norm(w)
        if beta[j - 1] < cut_off_tol:
            return V[:j], alpha[:j], beta[:j - 1]
        V[j] = w / beta[j - 1]
        w = A @ V[j]
        alpha[j] = w.dot(V[j])
        w = w - alpha[j] * V[j] - beta[j - 1] * V[j - 1]
        it = int(min(max(j - 1, 0), ortho_iters))
        for k in reversed(range(it)):
            w = w - V[k].dot(w) * V[k]
    return V, alpha, beta


def create_ritz_vec(A, rhs, num_ritz_vectors, ortho=True):
    """-> (ascending Ritz values, Ritz vectors [num, n])."""
    V, d, e = lanczos(A, rhs, num_ritz_vectors, np.inf if ortho else 0)
    vals, Q = scipy.linalg.eigh_tridiagonal(d, e, select="a")
    vecs = (V.T @ Q[:, :num_ritz_vectors]).T
    return vals, vecs


def training_rhs(ritz_vectors, coef, num_requested=None, cut_fraction=0.6, low_scale=9.0):
    """preprocess.py:38-58 with coef [num_ritz, count] given: rows b_i of the result are
    normalise(sum_k c'_{k,i} r_k), c' = coef with its first int(0.6 num_requested) rows x 9
    (num_requested: the number of Ritz vectors asked of Lanczos, preprocess.py:32,44 --
    more than len(ritz_vectors) when Lanczos stopped early)."""
    c = np.array(coef, dtype=np.float64, copy=True)
    cut = int((num_requested if num_requested is not None else len(ritz_vectors)) * cut_fraction)
    c[:cut] *= low_scale
    b = c.T @ ritz_vectors
    return b / np.linalg.norm(b, axis=1, keepdims=True)
