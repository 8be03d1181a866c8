"""FP64 restatement of the reference PCG-family loops (test infrastructure only).

All three follow cg_tests.py exactly in the order of operations that changes
floating-point results: the true residual r = b - A x is recomputed after every
update (cg_tests.py:57,83,165,260) and the stopping test is
||r|| / ||b|| > max(tol, atol / ||b||).
Vectors are numpy float64 arrays over the unknowns; A is a scipy sparse matrix;
predict(r) returns the preconditioned vector.
"""
from __future__ import annotations

import numpy as np


def npcg(b, A, x0, predict, max_it, tol=1e-10, atol=1e-12, trace=None):
    """cg_tests.py:26-101.  Returns (x, iterations, residual history).  trace: optional
    list that receives each search direction p (for property checks)."""
    norm_b = np.linalg.norm(b)
    r = b - A @ x0
    hist = [np.linalg.norm(r) / norm_b]
    x = x0.copy()
    z = predict(r)
    p = z
    if trace is not None:
        trace.append(p.copy())
    w = A @ p
    rz_old = r.dot(z)
    alpha = rz_old / p.dot(w)
    x = x + alpha * p
    r = b - A @ x
    norm = np.linalg.norm(r) / norm_b
    hist.append(norm)
    k = 1
    while norm > max(tol, atol / norm_b) and k < max_it:
        z = predict(r)
        rz = r.dot(z)
        beta = rz / rz_old
        p = z + beta * p
        if trace is not None:
            trace.append(p.copy())
        w = A @ p
        alpha = rz / p.dot(w)
        x = x + alpha * p
        r = b - A @ x
        k += 1
        rz_old = rz
        norm = np.linalg.norm(r) / norm_b
        hist.append(norm)
    return x, k, hist


def npcg_flex(b, A, x0, predict, max_it, tol=1e-10, atol=1e-12, trace=None):
    """cg_tests.py:103-183: beta = z.(r - r_old) / rz_old (Polak-Ribiere form).  trace:
    optional list that receives each search direction p."""
    norm_b = np.linalg.norm(b)
    r = b - A @ x0
    hist = [np.linalg.norm(r) / norm_b]
    x = x0.copy()
    z = predict(r)
    p = z
    if trace is not None:
        trace.append(p.copy())
    w = A @ p
    rz_old = r.dot(z)
    r_old = r
    alpha = rz_old / p.dot(w)
    x = x + alpha * p
    r = b - A @ x
    norm = np.linalg.norm(r) / norm_b
    hist.append(norm)
    k = 1
    while norm > max(tol, atol / norm_b) and k < max_it:
        z = predict(r)
        rz = r.dot(z)
        beta = z.dot(r - r_old) / rz_old
        p = z + beta * p
        if trace is not None:
            trace.append(p.copy())
        w = A @ p
        alpha = rz / p.dot(w)
        x = x + alpha * p
        r_old = r
        r = b - A @ x
        rz_old = rz
        k += 1
        norm = np.linalg.norm(r) / norm_b
        hist.append(norm)
    return x, k, hist


def npsd(b, A, x0, predict, max_it, tol=1e-10, atol=1e-10, num_prev=2, trace=None):
    """cg_tests.py:189-284: steepest descent A-orthogonalised against the last
    num_prev directions; stop when ||r||/||b|| < max(tol, atol/||b||).  trace: optional
    list that receives, per iteration, (new direction, the kept directions it was
    projected against, residual after the step)."""
    norm_b = np.linalg.norm(b)
    r = b - A @ x0
    hist = [np.linalg.norm(r) / norm_b]
    p, Ap, alfa = [], [], []
    x = x0.copy()
    norm = hist[0]
    for i in range(1, max_it + 1):
        q = predict(r)
        kept = [v.copy() for v in p] if trace is not None else None
        if num_prev > 0:
            if i > num_prev:
                for j in reversed(range(num_prev)):
                    q = q - q.dot(Ap[j]) / alfa[j] * p[j]
                for j in range(num_prev - 1):
                    p[j] = p[j + 1]
                    Ap[j] = Ap[j + 1]
                    alfa[j] = alfa[j + 1]
            else:
                for j in reversed(range(i - 1)):
                    q = q - q.dot(Ap[j]) / alfa[j] * p[j]
        if num_prev > 0:
            if i > num_prev:
                p[num_prev - 1] = q
                Ap[num_prev - 1] = A @ q
                alfa[num_prev - 1] = p[-1].dot(Ap[-1])
            else:
                p.append(q)
                Ap.append(A @ q)
                alfa.append(p[-1].dot(Ap[-1]))
            beta = p[-1].dot(r) / alfa[-1]
        else:
            alpha = q.dot(A @ q)
            beta = q.dot(r) / alpha
        x = x + q * beta
        r = b - A @ x
        if trace is not None:
            trace.append((q.copy(), kept, r.copy()))
        norm = np.linalg.norm(r) / norm_b
        hist.append(norm)
        if norm < max(tol, atol / norm_b):
            return x, i, hist
    return x, max_it, hist


def make_predict(model_fn, fluid_cells, shape, in_dtype=np.float32):
    """test.py:42-51: r -> normalize(r as float32) -> scatter to the grid ->
    model -> float64 -> gather fluid cells.  model_fn maps a (1,1,N1,N2,N3) float64
    torch tensor to the same shape."""
    import torch
    V = int(np.prod(shape))

    def predict(r):
        rf = r.astype(in_dtype).astype(np.float64)
        nrm = np.linalg.norm(rf)
        rn = rf / max(nrm, 1e-12)
        g = np.zeros(V)
        g[fluid_cells] = rn
        y = model_fn(torch.from_numpy(g.reshape((1, 1) + tuple(shape))))
        return y.reshape(-1).numpy()[fluid_cells].astype(np.float64)
    return predict
