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
from ._lib import lib, check, ptr, stream_ptr, code_of, MlpcgError

_KINDS = {
    "SmallSMModelDn3D": _lib.MODEL_SMALL_SM,
    "SPDSMModelDn3D": _lib.MODEL_SPD,
    "SmallSPDSMModelDn3D": _lib.MODEL_SMALL_SPD,
}


def _require_vec(t, what, dtype, numel, device):
    """The checks the C ABI cannot make (it takes raw pointers without lengths)."""
    if not isinstance(t, torch.Tensor):
        raise MlpcgError(f"{what} must be a torch tensor")
    if t.device != device:
        raise MlpcgError(f"{what} is on {t.device}, the plan on {device}")
    if t.dtype != dtype:
        raise MlpcgError(f"{what} has dtype {t.dtype}, expected {dtype}")
    if t.numel() != numel:
        raise MlpcgError(f"{what} has {t.numel()} elements, expected {numel}")
    if not t.is_contiguous():
        raise MlpcgError(f"{what} must be contiguous")
    if t.data_ptr() % 16:
        raise MlpcgError(f"{what} must be 16-byte aligned")


class Preconditioner:
    """z = P(r) for a fixed image.  dtype: storage type of K and of the level vectors
    (torch.float32 / bfloat16 / float16; float64 for checking).

    Execution policies of this plan (no effect on the result beyond FP32 summation order):
    fused_mode -- BF16 blocks on grids with n1, n2 % 8 == 0 and n3 % 64 == 0 (C <= 3)
      recompute their stencils on the matrix cores inside the apply instead of storing
      them: 0 never, 1 (default) on such grids of at least 2^18 voxels, 2 on every such grid;
    fold_third -- three-channel images whose third channel is the sum of the first two (the
      [air, fluid, non-solid] images of the reference's models): the fused blocks take two
      channels, the third channel's weights added to the others' (equal in exact arithmetic;
      the caller vouches for the relation, which the plan does not check);
    coarse_tail -- the two coarsest SmallSM levels run as one workgroup kernel when level
      n-1 is small enough (even dims; at most 4096 voxels with 16-bit storage and
      n3 % 4 == 0, else 2048)."""

    def __init__(self, model, image, dtype=torch.float32, fused_mode=1, coarse_tail=True, fold_third=False):
        kind = _KINDS.get(type(model).__name__)
        if kind is None:
            raise MlpcgError(f"no native plan for {type(model).__name__}")
        if not isinstance(image, torch.Tensor) or not image.is_cuda:
            raise MlpcgError("image must be a tensor on the GPU")
        if image.dim() != 4:
            raise MlpcgError(f"image must be (C, N1, N2, N3), got shape {tuple(image.shape)}")
        self.dtype = dtype
        self.acc = torch.float64 if dtype == torch.float64 else torch.float32
        self.vec_dtype = self.acc
        self.shape = tuple(int(s) for s in image.shape[-3:])
        self.numel = self.shape[0] * self.shape[1] * self.shape[2]
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
            raise MlpcgError(f"unsupported interpolation mode {mode}")
        handle = C.c_void_p()
        n1, n2, n3 = self.shape
        check(lib.mlpcg_precond_create_ex(C.byref(handle), kind, self.n, self.C, n1, n2, n3,
                                          code_of(dtype), int(mode == "nearest"),
                                          int(bool(getattr(model, "mask", False))),
                                          int(bool(getattr(model, "swap_ord", False))),
                                          int(fused_mode) | (4 if fold_third else 0),
                                          int(bool(coarse_tail)), ptr(self._img),
                                          warr, barr, stream_ptr(self.device)), "precond_create")
        self._h = handle
        self._level_numel = [(n1 >> l) * (n2 >> l) * (n3 >> l) for l in range(self.n + 1)]
        self._block_level = self._levels_of_blocks(kind)

    def _levels_of_blocks(self, kind):
        n = self.n
        if kind == _lib.MODEL_SMALL_SPD:
            return list(range(n)) + [n]
        return list(range(n)) + list(range(n)) + [n]

    def update(self, image, model=None):
        """Recompute pyramid, stencils and coefficients for a new image (same grid) and,
        if `model` is given, its current parameters; no allocation in the library."""
        if tuple(image.shape) != tuple(self._img.shape):
            raise MlpcgError(f"update: image shape {tuple(image.shape)} differs from the plan's "
                             f"{tuple(self._img.shape)}")
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

    @property
    def fused_blocks(self) -> int:
        """Blocks of this plan that run the fused matrix-core apply."""
        return int(lib.mlpcg_precond_fused_blocks(self._h))

    @property
    def coarse_tail(self) -> bool:
        return bool(lib.mlpcg_precond_coarse_tail(self._h) == 1)

    def _check_scale(self, in_scale):
        if in_scale is not None:
            _require_vec(in_scale, "in_scale", torch.float64, 1, self.device)

    def apply_block(self, k, x, out_dtype=None, in_scale=None):
        """y = block k (weights-array order) applied to x alone, on its level's grid; the
        fused matrix-core kernel when the plan uses it for that block."""
        if not 0 <= int(k) < len(self._block_level):
            raise MlpcgError(f"apply_block: block {k} out of range")
        out_dtype = out_dtype or x.dtype
        if x.dtype not in (self.dtype, self.vec_dtype) or out_dtype not in (self.dtype, self.vec_dtype):
            raise MlpcgError(f"apply_block: vector types {x.dtype} -> {out_dtype} do not fit a "
                             f"{self.dtype} plan")
        x = x.contiguous()
        _require_vec(x, "x", x.dtype, self._level_numel[self._block_level[int(k)]], self.device)
        self._check_scale(in_scale)
        y = torch.empty(x.shape, dtype=out_dtype, device=x.device)
        check(lib.mlpcg_precond_apply_block(self._h, int(k), ptr(x), code_of(x.dtype), ptr(y),
                                            code_of(out_dtype), ptr(in_scale),
                                            stream_ptr(self.device)), "precond_apply_block")
        return y

    def apply(self, r, in_scale=None, out_mask=None, out=None):
        """r: (..., N1, N2, N3) tensor of the vector type (float32; float64 for a float64
        plan).  in_scale: optional 0-dim float64 device tensor multiplying r."""
        if r.device != self.device:
            raise MlpcgError(f"apply: r is on {r.device}, the plan on {self.device}")
        r = r.to(self.vec_dtype).contiguous()
        _require_vec(r, "r", self.vec_dtype, self.numel, self.device)
        self._check_scale(in_scale)
        if out_mask is not None:
            _require_vec(out_mask, "out_mask", torch.uint8, self.numel, self.device)
        z = out if out is not None else torch.empty_like(r)
        _require_vec(z, "out", self.vec_dtype, self.numel, self.device)
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
