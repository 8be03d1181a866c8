"""Native preconditioner plan (image pyramid, per-block stencils K and SMLinear
coefficients computed once per image, then applied with one C call per application).

This is SmallSMModelDn3D.eval_forward with its imgs/c0/c1 caching
(sm_model_3d.py:180-262) extended to the stencil weights, run as a fixed kernel
sequence without host synchronisation.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from ._lib import lib, check, ptr, stream_ptr, code_of

_KINDS = {
    "SmallSMModelDn3D": _lib.MODEL_SMALL_SM,
    "SPDSMModelDn3D": _lib.MODEL_SPD,
    "SmallSPDSMModelDn3D": _lib.MODEL_SMALL_SPD,
}


def set_fused_apply(mode) -> int:
    """Process-wide policy for plans created afterwards: BF16 blocks on grids with
    n1, n2 % 8 == 0 and n3 % 64 == 0 recompute their stencils on the matrix cores inside
    the apply instead of storing them.  mode 0 / False: never; 1 / True (default): on such
    grids of at least 2^18 voxels; 2: on every such grid.  Returns the previous mode."""
    return int(lib.mlpcg_set_fused_apply(int(mode)))


def set_coarse_tail(enable) -> bool:
    """Process-wide switch (default on) for SmallSM plans created afterwards: the two
    coarsest levels run as one workgroup kernel when level n-1 is small enough (even
    dims; at most 4096 voxels with 16-bit storage and n3 % 4 == 0, else 2048).
    Returns the previous setting."""
    return bool(lib.mlpcg_set_coarse_tail(int(bool(enable))))


class Preconditioner:
    """z = P(r) for a fixed image.  dtype: storage type of K and of the level vectors
    (torch.float32 / bfloat16 / float16; float64 for checking)."""

    def __init__(self, model, image, dtype=torch.float32):
        kind = _KINDS.get(type(model).__name__)
        if kind is None:
            raise _lib.MlpcgError(f"no native plan for {type(model).__name__}")
        if not image.is_cuda:
            raise _lib.MlpcgError("image must be on the GPU")
        self.dtype = dtype
        self.acc = torch.float64 if dtype == torch.float64 else torch.float32
        self.vec_dtype = self.acc
        self.shape = tuple(int(s) for s in image.shape[-3:])
        self.C = int(image.shape[0])
        self.n = int(model.n)
        self.device = image.device
        ws, bs = model.plan_params()
        # keep the converted parameters alive for the plan's setup
        self._w = [w.detach().to(self.device, self.acc).contiguous() for w in ws]
        self._b = [b.detach().to(self.device, self.acc).contiguous() for b in bs]
        self._img = image.detach().to(self.acc).contiguous()
        warr = (C.c_void_p * len(self._w))(*[ptr(w) for w in self._w])
        barr = (C.c_void_p * len(self._b))(*[ptr(b) for b in self._b])
        mode = getattr(model, "mode", "nearest")
        if mode not in ("trilinear", "nearest"):
            raise _lib.MlpcgError(f"unsupported interpolation mode {mode}")
        handle = C.c_void_p()
        n1, n2, n3 = self.shape
        check(lib.mlpcg_precond_create(C.byref(handle), kind, self.n, self.C, n1, n2, n3,
                                       code_of(dtype), int(mode == "nearest"),
                                       int(bool(getattr(model, "mask", False))),
                                       int(bool(getattr(model, "swap_ord", False))), ptr(self._img),
                                       warr, barr, stream_ptr(self.device)), "precond_create")
        self._h = handle

    def update(self, image, model=None):
        """Recompute pyramid, stencils and coefficients for a new image (same grid) and,
        if `model` is given, its current parameters; no allocation in the library."""
        if model is not None:
            ws, bs = model.plan_params()
            for dst, src in zip(self._w, ws):
                dst.copy_(src.detach())
            for dst, src in zip(self._b, bs):
                dst.copy_(src.detach())
        self._img.copy_(image.detach())
        warr = (C.c_void_p * len(self._w))(*[ptr(w) for w in self._w])
        barr = (C.c_void_p * len(self._b))(*[ptr(b) for b in self._b])
        check(lib.mlpcg_precond_update(self._h, ptr(self._img), warr, barr, stream_ptr(self.device)),
              "precond_update")
        return self

    @property
    def handle(self):
        return self._h

    @property
    def nbytes(self) -> int:
        return int(lib.mlpcg_precond_bytes(self._h))

    def apply_block(self, k, x, out_dtype=None, in_scale=None):
        """y = block k (weights-array order) applied to x alone, on its level's grid; the
        fused matrix-core kernel when the plan uses it for that block."""
        out_dtype = out_dtype or x.dtype
        x = x.contiguous()
        y = torch.empty(x.shape, dtype=out_dtype, device=x.device)
        check(lib.mlpcg_precond_apply_block(self._h, int(k), ptr(x), code_of(x.dtype), ptr(y),
                                            code_of(out_dtype), ptr(in_scale),
                                            stream_ptr(self.device)), "precond_apply_block")
        return y

    def apply(self, r, in_scale=None, out_mask=None, out=None):
        """r: (..., N1, N2, N3) tensor of the vector type (float32; float64 for a float64
        plan).  in_scale: optional 0-dim float64 device tensor multiplying r."""
        r = r.to(self.vec_dtype).contiguous()
        z = out if out is not None else torch.empty_like(r)
        check(lib.mlpcg_precond_apply(self._h, ptr(r), ptr(z), ptr(in_scale), ptr(out_mask),
                                      stream_ptr(self.device)), "precond_apply")
        return z

    __call__ = apply

    def close(self):
        if getattr(self, "_h", None):
            lib.mlpcg_precond_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
