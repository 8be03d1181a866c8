"""Ritz-vector training right-hand sides on the GPU (lib/create_dataset.py:28-70 and
preprocess.py:38-58).

The reference builds its training data offline: a Lanczos process from the scene's
right-hand side with full reorthogonalisation (numpy, FP64), the Ritz vectors of the
tridiagonal matrix (scipy eigh_tridiagonal), then right-hand sides as random combinations
of the Ritz vectors with the low-eigenvalue 60 % weighted 9x.  Here the Lanczos vectors
live in HBM in FP64: the matrix product is the library's CSR FP64 SpMV, the dots, the
three-term update and the normalisation are the library's FP64 kernels with the scalars
kept on the device, the reorthogonalisation against the stored basis is two classical
Gram-Schmidt passes (GEMV pairs on rocBLAS; the same projection as the reference's
one-vector-at-a-time loop in exact arithmetic, and as orthogonal in floating point), and
the Ritz vectors and right-hand sides are GEMMs.  The 1600 x 1600 tridiagonal
eigenproblem stays on the host (scipy, as in the reference).
"""
from __future__ import annotations

import numpy as np
import scipy.linalg
import torch

from . import _lib
from ._lib import lib, check, ptr, stream_ptr, MlpcgError

_ALIGN = 8  # doubles: rows of the basis start on 64-byte boundaries


class CSRMatrix64:
    """The fluid Laplacian in CSR with FP64 values on the GPU (int32 indices)."""

    def __init__(self, A, device="cuda"):
        A = A.tocsr()
        self.shape = A.shape
        self.device = torch.device(device)
        self.rowptr = torch.as_tensor(A.indptr.astype(np.int32), device=self.device)
        self.col = torch.as_tensor(A.indices.astype(np.int32), device=self.device)
        self.val = torch.as_tensor(A.data.astype(np.float64), device=self.device)

    def matvec(self, x, out):
        check(lib.mlpcg_spmv_csr(ptr(self.rowptr), ptr(self.col), ptr(self.val), ptr(x), ptr(out),
                                 int(self.shape[0]), _lib.F64, stream_ptr(self.device)), "spmv_csr f64")
        return out


class _Vec:
    """FP64 vector kernels with the scalars on the device."""

    def __init__(self, device):
        self.device = device
        self.ws = torch.empty(max(16, int(lib.mlpcg_dot_f64_workspace_size())), dtype=torch.uint8,
                              device=device)

    def dot(self, a, b, out):
        check(lib.mlpcg_dot_f64(ptr(a), ptr(b), a.numel(), ptr(out), ptr(self.ws), stream_ptr(self.device)),
              "dot_f64")

    def update(self, w, v1, v0, alpha, beta):
        check(lib.mlpcg_lanczos_update(ptr(w), ptr(v1), ptr(v0), ptr(alpha), ptr(beta), w.numel(),
                                       stream_ptr(self.device)), "lanczos_update")

    def normalize(self, out, w, nrm2, norm_out):
        check(lib.mlpcg_normalize_f64(ptr(out), ptr(w), ptr(nrm2), ptr(norm_out), w.numel(),
                                      stream_ptr(self.device)), "normalize_f64")


def _basis(m, n, device):
    """[m, n] FP64 view whose rows start on 64-byte boundaries (16-byte vector loads)."""
    npad = (n + _ALIGN - 1) // _ALIGN * _ALIGN
    return torch.zeros(m, npad, dtype=torch.float64, device=device)[:, :n]


def lanczos(A: CSRMatrix64, init_v, m: int, ortho: bool = True, cut_off_tol: float = 1e-10):
    """lib/create_dataset.py:28-52.  -> (V [k, n] device FP64 basis rows, alpha [k] and
    beta [k-1] host FP64), k = m unless beta fell below cut_off_tol (then k = j, as in the
    reference).  init_v: host or device vector over the unknowns."""
    dev = A.device
    n = int(A.shape[0])
    if m < 1:
        raise MlpcgError("lanczos: m must be >= 1")
    v0 = torch.as_tensor(np.asarray(init_v, dtype=np.float64) if not torch.is_tensor(init_v) else init_v,
                         dtype=torch.float64, device=dev).reshape(-1)
    if v0.numel() != n:
        raise MlpcgError(f"lanczos: start vector has {v0.numel()} entries, A has {n} rows")
    V = _basis(m, n, dev)
    ops = _Vec(dev)
    sc = torch.zeros(3 * m + 2, dtype=torch.float64, device=dev)  # alpha[m], beta[m], norms^2[m+2]
    alpha, beta, nrm2 = sc[:m], sc[m:2 * m], sc[2 * m:]
    w = _basis(1, n, dev)[0]
    tmp = _basis(1, n, dev)[0]
    u = v0.contiguous() if v0.data_ptr() % 16 == 0 else _basis(1, n, dev)[0].copy_(v0)
    ops.dot(u, u, nrm2[0:1])
    ops.normalize(V[0], u, nrm2[0:1], None)
    A.matvec(V[0], w)
    ops.dot(w, V[0], alpha[0:1])
    ops.update(w, V[0], None, alpha[0:1], None)
    k = m
    for j in range(1, m):
        ops.dot(w, w, nrm2[j:j + 1])
        ops.normalize(V[j], w, nrm2[j:j + 1], beta[j - 1:j])
        if float(beta[j - 1]) < cut_off_tol:  # one host read per step, as the reference's test
            k = j
            break
        A.matvec(V[j], w)
        ops.dot(w, V[j], alpha[j:j + 1])
        ops.update(w, V[j], V[j - 1], alpha[j:j + 1], beta[j - 1:j])
        if ortho and j >= 2:  # against v_0 .. v_{j-2}, like create_dataset.py:50-52
            B = V[:j - 1]
            for _ in range(2):
                h = torch.mv(B, w)
                torch.addmv(w, B.t(), h, alpha=-1.0, out=tmp)
                w.copy_(tmp)
    a = alpha[:k].cpu().numpy().copy()
    b = beta[:max(k - 1, 0)].cpu().numpy().copy()
    return V[:k], a, b


def ritz_vectors(A: CSRMatrix64, rhs, num_ritz_vectors: int, ortho: bool = True):
    """lib/create_dataset.py:54-68.  -> (ascending Ritz values [k] host, Ritz vectors
    [k, n] device FP64)."""
    V, d, e = lanczos(A, rhs, num_ritz_vectors, ortho)
    vals, Q = scipy.linalg.eigh_tridiagonal(d, e, select="a")
    Qd = torch.as_tensor(Q[:, :num_ritz_vectors], dtype=torch.float64, device=V.device)
    R = _basis(Qd.shape[1], V.shape[1], V.device)
    R.copy_(Qd.t() @ V)
    return vals, R


def training_rhs(R, count=None, coef=None, num_requested=None, generator=None, dtype=torch.float32):
    """preprocess.py:38-58: rows b_i = normalise(sum_k c_ki w_k r_k) with c ~ N(0, 1) and
    w_k = 9 for the first int(0.6 num_requested) Ritz vectors (the smallest Ritz values),
    1 for the rest.  coef [k, count] may be given (e.g. to match a host computation);
    otherwise it is drawn on the device from `generator`.  -> [count, n] in `dtype`."""
    k = R.shape[0]
    if coef is None:
        if count is None:
            raise MlpcgError("training_rhs: give count or coef")
        coef = torch.randn(k, int(count), generator=generator, dtype=torch.float64, device=R.device)
    else:
        coef = torch.as_tensor(coef, dtype=torch.float64, device=R.device).clone()
    if coef.shape[0] != k:
        raise MlpcgError(f"training_rhs: coef has {coef.shape[0]} rows, there are {k} Ritz vectors")
    cut = int((num_requested if num_requested is not None else k) * 0.6)
    coef[:cut] *= 9.0
    B = coef.t() @ R
    ops = _Vec(R.device)
    out = _basis(B.shape[0], B.shape[1], R.device)
    out.copy_(B)
    nrm2 = torch.zeros(B.shape[0], dtype=torch.float64, device=R.device)
    for i in range(B.shape[0]):
        ops.dot(out[i], out[i], nrm2[i:i + 1])
        ops.normalize(out[i], out[i], nrm2[i:i + 1], None)
    return out.to(dtype)


def full_grid(b, fluid_idx, V):
    """Right-hand sides over the fluid unknowns -> full-grid vectors (zero elsewhere)."""
    out = torch.zeros(b.shape[0], V, dtype=b.dtype, device=b.device)
    out[:, fluid_idx] = b
    return out
