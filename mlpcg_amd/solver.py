"""PCG solvers on the GPU.

* FusedPCG: the native driver (libmlpcg_hip.so) -- the whole npcg / npcg_flex / npsd
  loop runs as a fixed kernel sequence with the scalars on the device; the host only
  watches a status flag a few iterations behind.  Matrix-free on the full grid
  ("flags" layout) or CSR over the fluid unknowns ("csr" layout).
* CSRMatrix: a drop-in for the torch sparse CSR `A` of cg_tests.py (`A @ p`).
* npcg / npcg_flex / npsd: the reference loops (cg_tests.py:26-284) with the same
  signatures and return values (x, k, timer, norm) on torch tensors.
* model_predict: test.py:42-51.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib, kernels
from ._lib import lib, check, ptr, stream_ptr
from .precond import _require_vec
from .scenes import FLUID
from .timer import GlobalClock

METHODS = {"npcg": _lib.NPCG, "npcg_flex": _lib.NPCG_FLEX, "npsd": _lib.NPSD}


class CSRMatrix:
    """Fluid Laplacian in CSR (int32 indices) on the GPU; `A @ x` runs the HIP SpMV."""

    def __init__(self, A, device="cuda", dtype=torch.float32):
        A = A.tocsr()
        self.shape = A.shape
        self.dtype = dtype
        self.rowptr = torch.as_tensor(A.indptr.astype(np.int32), device=device)
        self.col = torch.as_tensor(A.indices.astype(np.int32), device=device)
        self.val = torch.as_tensor(A.data, dtype=dtype, device=device)
        self.nnz = int(A.nnz)

    def __matmul__(self, x):
        x = x.to(self.dtype).contiguous()
        return kernels.spmv_csr(self.rowptr, self.col, self.val, x)

    def to(self, dtype):
        out = object.__new__(CSRMatrix)
        out.__dict__.update(self.__dict__)
        out.dtype = dtype
        out.val = self.val.to(dtype)
        return out


class FusedPCG:
    """Native PCG for one scene.

    flags: (N1, N2, N3) integer flag codes (lib/read_data.py:13-16).
    layout: "flags" (full-grid vectors, matrix-free A) or "csr" (vectors over the fluid
    cells, A given as a scipy CSR matrix in fluid order).
    """

    def __init__(self, flags, layout="flags", A=None, device="cuda", graphs=False):
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise _lib.MlpcgError(f"FusedPCG runs on the GPU, not on {self.device}")
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        flags_np = np.asarray(flags)
        if flags_np.ndim != 3:
            raise _lib.MlpcgError(f"flags must be a 3-d grid, got shape {flags_np.shape}")
        self.shape = tuple(int(s) for s in flags_np.shape)
        self.layout = layout
        self._flags = torch.as_tensor(flags_np.astype(np.int8), device=self.device).contiguous()
        n1, n2, n3 = self.shape
        h = C.c_void_p()
        st = stream_ptr(self.device)
        if layout == "flags":
            check(lib.mlpcg_solver_create_flags(C.byref(h), ptr(self._flags), n1, n2, n3, st),
                  "solver_create_flags")
            self.n = n1 * n2 * n3
        elif layout == "csr":
            if A is None:
                raise _lib.MlpcgError("csr layout needs the matrix A")
            self.A = A if isinstance(A, CSRMatrix) else CSRMatrix(A, self.device)
            fluid = np.flatnonzero(flags_np.ravel() == FLUID).astype(np.int32)
            self._fluid = torch.as_tensor(fluid, device=self.device)
            self.n = int(fluid.size)
            if self.A.shape[0] != self.n:
                raise _lib.MlpcgError("A does not match the fluid cells of flags")
            check(lib.mlpcg_solver_create_csr(C.byref(h), ptr(self.A.rowptr), ptr(self.A.col),
                                              ptr(self.A.val), self.n, ptr(self._fluid), n1, n2,
                                              n3, st), "solver_create_csr")
        else:
            raise ValueError(f"unknown layout {layout}")
        self._h = h
        self.set_graphs(graphs)

    def set_graphs(self, enable: bool):
        """Replay each iteration as a captured HIP graph (bit-identical results)."""
        check(lib.mlpcg_solver_set_graphs(self._h, int(bool(enable))), "solver_set_graphs")
        self.graphs = bool(enable)

    def vector(self, values):
        """Device float32 vector in this solver's layout from fluid-cell values."""
        v = torch.as_tensor(np.asarray(values), dtype=torch.float32)
        if self.layout == "csr":
            return v.to(self.device).contiguous()
        full = torch.zeros(self.shape[0] * self.shape[1] * self.shape[2], dtype=torch.float32)
        fluid = torch.as_tensor(np.flatnonzero(self._flags.cpu().numpy().ravel() == FLUID))
        full[fluid] = v
        return full.to(self.device)

    def solve(self, b, precond=None, x0=None, method="npcg", max_it=100, tol=1e-6, atol=1e-12,
              history=False, out=None):
        """-> (x, info dict).  b: float32 device vector in this layout."""
        if method not in METHODS:
            raise _lib.MlpcgError(f"unknown method {method}")
        if int(max_it) < 1:
            raise _lib.MlpcgError("max_it must be >= 1")
        if not isinstance(b, torch.Tensor) or b.device != self.device:
            raise _lib.MlpcgError(f"b must be a tensor on {self.device}")
        b = b.to(torch.float32).contiguous()
        _require_vec(b, "b", torch.float32, self.n, self.device)
        if x0 is not None:
            _require_vec(x0, "x0", torch.float32, self.n, self.device)
        x = out if out is not None else torch.empty_like(b)
        _require_vec(x, "out", torch.float32, self.n, self.device)
        if precond is not None:
            if precond.device != self.device or precond.shape != self.shape:
                raise _lib.MlpcgError(f"preconditioner grid {precond.shape} on {precond.device} does "
                                      f"not match the solver grid {self.shape} on {self.device}")
            if precond.dtype == torch.float64:
                raise _lib.MlpcgError("FP64 plans are for checking only; use float32/bfloat16/float16")
        hist = np.zeros(int(max_it) + 1, dtype=np.float64) if history else None
        info = _lib.SolveInfo()
        P = precond.handle if precond is not None else None
        check(lib.mlpcg_solver_solve(self._h, P, ptr(b), ptr(x0), ptr(x), METHODS[method], int(max_it),
                                     float(tol), float(atol),
                                     hist.ctypes.data if hist is not None else None, C.byref(info),
                                     stream_ptr(self.device)), "solver_solve")
        res = {"iters": info.iters, "status": info.status, "rel_res": info.rel_res,
               "norm_b": info.norm_b}
        if hist is not None:
            res["history"] = hist[: info.iters + 1].copy()
        return x, res

    def close(self):
        if getattr(self, "_h", None):
            lib.mlpcg_solver_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Reference-signature loops on torch tensors (cg_tests.py:26-284)
# ---------------------------------------------------------------------------

def _dot(a, b):
    if a.is_cuda and a.dtype == torch.float32 and b.dtype == torch.float32:
        return kernels.dot(a.contiguous(), b.contiguous())
    return a.dot(b)


def _norm(a):
    return float(torch.sqrt(_dot(a, a)))


def npcg(b, A, x_init, model_predict, max_it, tol=1e-10, atol=1e-12, verbose=False, callback=None):
    """cg_tests.py:26-101."""
    timer = GlobalClock()
    timer.start("Total")
    norm_b = _norm(b)
    r = b - A @ x_init
    norm = _norm(r) / norm_b
    if verbose:
        print(f"Iter 0, residual {norm}")
    x = torch.clone(x_init)
    z = model_predict(r, timer)
    p = z
    w = A @ p
    rz = rz_old = _dot(r, z)
    alpha = rz / _dot(p, w)
    x += (alpha * p).to(x.dtype)
    r = b - A @ x
    norm = _norm(r) / norm_b
    if callback:
        callback(norm)
    k = 1
    while norm > max(tol, atol / norm_b) and k < max_it:
        z = model_predict(r, timer)
        rz = _dot(r, z)
        beta = rz / rz_old
        p = z + (beta * p).to(z.dtype)
        w = A @ p
        alpha = rz / _dot(p, w)
        x += (alpha * p).to(x.dtype)
        r = b - A @ x
        k += 1
        rz_old = rz
        norm = _norm(r) / norm_b
        if callback:
            callback(norm)
        if verbose:
            print(f"Iter {k}, residual {norm}")
    timer.stop("Total")
    return x, k, timer, norm


def npcg_flex(b, A, x_init, model_predict, max_it, tol=1e-10, atol=1e-12, verbose=False, callback=None):
    """cg_tests.py:103-183."""
    timer = GlobalClock()
    timer.start("Total")
    norm_b = _norm(b)
    r = b - A @ x_init
    x = torch.clone(x_init)
    z = model_predict(r, timer)
    p = z
    w = A @ p
    rz = rz_old = _dot(r, z)
    r_old = r
    alpha = rz / _dot(p, w)
    x += (alpha * p).to(x.dtype)
    r = b - A @ x
    norm = _norm(r) / norm_b
    if callback:
        callback(norm)
    k = 1
    while norm > max(tol, atol / norm_b) and k < max_it:
        z = model_predict(r, timer)
        rz = _dot(r, z)
        beta = _dot(z, r - r_old) / rz_old
        p = z + (beta * p).to(z.dtype)
        w = A @ p
        alpha = rz / _dot(p, w)
        x += (alpha * p).to(x.dtype)
        r_old = r
        r = b - A @ x
        rz_old = rz
        k += 1
        norm = _norm(r) / norm_b
        if callback:
            callback(norm)
        if verbose:
            print(f"Iter {k}, residual {norm}")
    timer.stop("Total")
    return x, k, timer, norm


def npsd(b, A, x_init, model_predict, max_it, tol=1e-10, atol=1e-10, verbose=False, callback=None):
    """cg_tests.py:189-284 (num_prev = 2)."""
    timer = GlobalClock()
    timer.start("Total")
    norm_b = _norm(b)
    r = b - A @ x_init
    num_prev = 2
    p, Ap, alfa = [], [], []
    x_sol = torch.clone(x_init)
    imgs, c0, c1 = [], [], []
    norm = _norm(r) / norm_b
    for i in range(1, max_it + 1):
        q = model_predict(r, timer, imgs, c0, c1)
        if i > num_prev:
            for j in reversed(range(num_prev)):
                q = q - (_dot(q, Ap[j]) / alfa[j] * p[j]).to(q.dtype)
            for j in range(num_prev - 1):
                p[j], Ap[j], alfa[j] = p[j + 1], Ap[j + 1], alfa[j + 1]
            p[num_prev - 1] = q
            Ap[num_prev - 1] = A @ q
            alfa[num_prev - 1] = _dot(p[-1], Ap[-1])
        else:
            for j in reversed(range(i - 1)):
                q = q - (_dot(q, Ap[j]) / alfa[j] * p[j]).to(q.dtype)
            p.append(q)
            Ap.append(A @ q)
            alfa.append(_dot(p[-1], Ap[-1]))
        beta = _dot(p[-1], r) / alfa[-1]
        x_sol += (q * beta).to(x_sol.dtype)
        r = b - A @ x_sol
        norm = _norm(r) / norm_b
        if callback:
            callback(norm)
        if verbose:
            print(f"Iter {i}, residual {norm}")
        if norm < max(tol, atol / norm_b):
            timer.stop("Total")
            return x_sol, i, timer, norm
    timer.stop("Total")
    return x_sol, max_it, timer, norm


def model_predict(model, image, fluid_cells, dtype=torch.float32):
    """test.py:42-51: normalise r, scatter it onto the grid, run eval_forward, gather."""
    shape = tuple(image.shape[1:])

    def predict(r, timer=None, imgs=None, c0=None, c1=None):
        imgs = [] if imgs is None else imgs
        c0 = [] if c0 is None else c0
        c1 = [] if c1 is None else c1
        out_dtype = r.dtype
        with torch.no_grad():
            r = r.to(dtype)
            r = r / torch.clamp(torch.sqrt(_dot(r.float(), r.float())).to(dtype), min=1e-12)
            b = torch.zeros(int(np.prod(shape)), device=r.device, dtype=dtype)
            b[fluid_cells] = r
            x = model.eval_forward(image, b.view((1, 1) + shape), timer, imgs, c0, c1).flatten()
        return x[fluid_cells].to(out_dtype)
    return predict
