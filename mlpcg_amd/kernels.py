"""Tensor-level wrappers of the C ABI (one function per entry point family).

Shapes follow the reference: image (C, N1, N2, N3); x (B, 1, N1, N2, N3) or any
tensor whose last three dims are the grid.  Weights are (27, C, 3, 3, 3) + (27,) for SM
blocks and (1, C, 3, 3, 3) + (1,) for SM linear, passed in the accumulate type (float64
for float64 data, float32 otherwise).  All work is queued on torch's current stream.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import weakref

import torch

from . import _lib
from ._lib import lib, check, ptr, stream_ptr, code_of, require_cuda, ACC_DTYPE


def _grid(t):
    n1, n2, n3 = t.shape[-3:]
    return int(n1), int(n2), int(n3)


def _batch(x):
    n1, n2, n3 = _grid(x)
    return x.numel() // (n1 * n2 * n3)


def _acc(dtype):
    return ACC_DTYPE[code_of(dtype)]


def _params(weight, bias, dtype):
    acc = _acc(dtype)
    return weight.detach().to(acc).contiguous(), bias.detach().to(acc).contiguous()


def workspace(nbytes, device):
    return torch.empty(max(int(nbytes), 16), dtype=torch.uint8, device=device)


# ------------------------------------------------------------------ SM block / transpose

def _block_forward(fn, image, x, weight, bias):
    require_cuda(image, x)
    dt = code_of(x.dtype)
    image = image.to(x.dtype).contiguous()
    W, b = _params(weight, bias, x.dtype)
    n1, n2, n3 = _grid(x)
    B = _batch(x)
    Cimg = image.shape[0]
    y = torch.empty_like(x)
    ws = workspace(lib.mlpcg_sm_block_workspace_size(Cimg, B, n1, n2, n3, dt), x.device)
    check(fn(ptr(image), ptr(x), ptr(W), ptr(b), ptr(y), ptr(ws), Cimg, B, n1, n2, n3, dt,
             stream_ptr(x.device)), fn.__name__)
    return y


def _block_backward(fn, g, image, x, weight, bias):
    require_cuda(g, image, x)
    dt = code_of(x.dtype)
    g = g.to(x.dtype).contiguous()
    image = image.to(x.dtype).contiguous()
    W, b = _params(weight, bias, x.dtype)
    n1, n2, n3 = _grid(x)
    B = _batch(x)
    Cimg = image.shape[0]
    gx = torch.empty_like(x)
    gW = torch.empty_like(W)
    gb = torch.empty_like(b)
    ws = workspace(lib.mlpcg_sm_block_workspace_size(Cimg, B, n1, n2, n3, dt), x.device)
    check(fn(ptr(g), ptr(image), ptr(x), ptr(W), ptr(b), ptr(gx), ptr(gW), ptr(gb), ptr(ws), Cimg,
             B, n1, n2, n3, dt, stream_ptr(x.device)), fn.__name__)
    return gx, gW.to(weight.dtype), gb.to(bias.dtype)


# FP16 copies of BF16 images kept by callers that reuse an image over many calls (a
# training scene's pyramid): the block entry points then skip converting the image for
# the matrix-core path (mlpcg_sm_block_*_ex).  id(image) -> (weak reference, copy); an
# entry goes when its image does.
_F16_COPIES = {}


def register_f16_image(image):
    """Keep an FP16 copy of the BF16 image `image` for the BF16 block calls that take it
    (the conversion the matrix-core path would otherwise make per call); returns it."""
    if image.dtype != torch.bfloat16:
        raise ValueError("register_f16_image: the image must be bfloat16")
    copy = image.contiguous().to(torch.float16)
    key = id(image)
    _F16_COPIES[key] = (weakref.ref(image, lambda _, k=key: _F16_COPIES.pop(k, None)), copy)
    return copy


def _f16_copy(image, x):
    if x.dtype != torch.bfloat16 or image.dtype != torch.bfloat16:
        return None
    e = _F16_COPIES.get(id(image))
    if e is None or e[0]() is not image:
        return None
    c = e[1]
    return c if c.shape == image.shape else None


def fused_pack(weight, bias):
    """The weights and bias of a BF16 block packed for the matrix-core path (uint8 device
    tensor of mlpcg_fused_pack_bytes(C) bytes)."""
    require_cuda(weight, bias)
    W, b = weight.detach().float().contiguous(), bias.detach().float().contiguous()
    Cimg = int(W.shape[1])
    out = torch.empty(lib.mlpcg_fused_pack_bytes(Cimg), dtype=torch.uint8, device=W.device)
    check(lib.mlpcg_fused_pack(ptr(W), ptr(b), Cimg, ptr(out), stream_ptr(W.device)), "fused_pack")
    return out


# Packings made once for a stretch of calls in which the weights do not change (one
# training step over several scenes), keyed by the (weight, bias) storage.
_PACKS = {}


@contextlib.contextmanager
def packed_weights(pairs):
    """Inside the block, BF16 block calls on these (weight, bias) pairs -- or on tensors
    sharing their storage -- use one packing made on entry; the weights must not change
    until the block ends."""
    keys = []
    try:
        for weight, bias in pairs:
            W, b = weight.detach().float().contiguous(), bias.detach().float().contiguous()
            key = (W.data_ptr(), b.data_ptr())
            _PACKS[key] = fused_pack(W, b)
            keys.append(key)
        yield
    finally:
        for key in keys:
            _PACKS.pop(key, None)


def _ex_args(image, x, weight, bias):
    """(image16, wpack) for the _ex block entry points, or None when neither is kept."""
    if x.dtype != torch.bfloat16:
        return None
    c16 = _f16_copy(image, x)
    pk = _PACKS.get((weight.data_ptr(), bias.data_ptr())) if weight.dtype == torch.float32 else None
    if c16 is None and pk is None:
        return None
    return (ptr(c16) if c16 is not None else None, ptr(pk) if pk is not None else None)


def sm_block_forward(image, x, weight, bias):
    ex = _ex_args(image, x, weight, bias)
    if ex is not None:
        fn = lambda im, *a: lib.mlpcg_sm_block_forward_ex(im, ex[0], ex[1], *a)
        fn.__name__ = "mlpcg_sm_block_forward_ex"
        return _block_forward(fn, image, x, weight, bias)
    return _block_forward(lib.mlpcg_sm_block_forward, image, x, weight, bias)


def sm_block_inference(image, x, weight, bias):
    return _block_forward(lib.mlpcg_sm_block_inference, image, x, weight, bias)


def sm_block_backward(g, image, x, weight, bias):
    ex = _ex_args(image, x, weight, bias)
    if ex is not None:
        fn = lambda gp, im, *a: lib.mlpcg_sm_block_backward_ex(gp, im, ex[0], ex[1], *a)
        fn.__name__ = "mlpcg_sm_block_backward_ex"
        return _block_backward(fn, g, image, x, weight, bias)
    return _block_backward(lib.mlpcg_sm_block_backward, g, image, x, weight, bias)


def sm_block_trans_forward(image, x, weight, bias):
    return _block_forward(lib.mlpcg_sm_block_trans_forward, image, x, weight, bias)


def sm_block_trans_inference(image, x, weight, bias):
    return _block_forward(lib.mlpcg_sm_block_trans_inference, image, x, weight, bias)


def sm_block_trans_backward(g, image, x, weight, bias):
    return _block_backward(lib.mlpcg_sm_block_trans_backward, g, image, x, weight, bias)


# ------------------------------------------------------------------ SM linear

def sm_linear(image, weight, bias):
    """-> (z of shape (1,), ybar of shape (1, C, 3, 3, 3)), accumulate type."""
    require_cuda(image)
    dt = code_of(image.dtype)
    acc = _acc(image.dtype)
    image = image.contiguous()
    w, b = _params(weight, bias, image.dtype)
    Cimg = image.shape[0]
    n1, n2, n3 = _grid(image)
    z = torch.empty(1, dtype=acc, device=image.device)
    ybar = torch.empty(1, Cimg, 3, 3, 3, dtype=acc, device=image.device)
    ws = workspace(lib.mlpcg_sm_linear_workspace_size(Cimg, n1, n2, n3, dt), image.device)
    check(lib.mlpcg_sm_linear_forward(ptr(image), ptr(w), ptr(b), ptr(z), ptr(ybar), ptr(ws), Cimg,
                                      n1, n2, n3, dt, stream_ptr(image.device)), "sm_linear")
    return z, ybar


def sm_linear_backward(g, ybar):
    require_cuda(g, ybar)
    dt = code_of(ybar.dtype)
    g = g.to(ybar.dtype).reshape(1).contiguous()
    gw = torch.empty_like(ybar)
    gb = torch.empty(1, dtype=ybar.dtype, device=ybar.device)
    check(lib.mlpcg_sm_linear_backward(ptr(g), ptr(ybar), ptr(gw), ptr(gb), ybar.shape[1], dt,
                                       stream_ptr(ybar.device)), "sm_linear_backward")
    return gw, gb


# ------------------------------------------------------------------ streamed-stencil parts

def stencil_weights(image, weight, bias, dtype=None, transposed=False):
    """K (27, N1, N2, N3) in `dtype` (default: image dtype)."""
    require_cuda(image)
    dtype = dtype or image.dtype
    image = image.contiguous()
    if image.dtype != dtype and image.dtype != torch.float32:
        image = image.to(dtype)
    W, b = _params(weight, bias, dtype)
    n1, n2, n3 = _grid(image)
    K = torch.empty(27, n1, n2, n3, dtype=dtype, device=image.device)
    check(lib.mlpcg_stencil_weights(ptr(image), code_of(image.dtype), ptr(W), ptr(b), ptr(K),
                                    code_of(dtype), int(transposed), image.shape[0], n1, n2, n3,
                                    stream_ptr(image.device)), "stencil_weights")
    return K


def sm_apply(K, x, out_dtype=None, prev=None, c0=None, c1=None, mask_a=None, mask_b=None,
             in_scale=None):
    require_cuda(K, x, prev, c0, c1, mask_a, mask_b, in_scale)
    out_dtype = out_dtype or x.dtype
    acc = _acc(K.dtype)
    c0 = None if c0 is None else c0.to(acc).contiguous()
    c1 = None if c1 is None else c1.to(acc).contiguous()
    in_scale = None if in_scale is None else in_scale.to(torch.float64).contiguous()
    mask_a = None if mask_a is None else mask_a.to(torch.uint8).contiguous()
    mask_b = None if mask_b is None else mask_b.to(torch.uint8).contiguous()
    n1, n2, n3 = _grid(x)
    y = torch.empty(x.shape, dtype=out_dtype, device=x.device)
    check(lib.mlpcg_sm_apply(ptr(K), code_of(K.dtype), ptr(x), code_of(x.dtype), ptr(y),
                             code_of(out_dtype), ptr(prev), ptr(c0), ptr(c1), ptr(mask_a),
                             ptr(mask_b), ptr(in_scale), _batch(x), n1, n2, n3,
                             stream_ptr(x.device)), "sm_apply")
    return y


def avgpool2(x):
    require_cuda(x)
    n1, n2, n3 = _grid(x)
    y = torch.empty(*x.shape[:-3], n1 // 2, n2 // 2, n3 // 2, dtype=x.dtype, device=x.device)
    check(lib.mlpcg_avgpool2(ptr(x), ptr(y), _batch(x), n1, n2, n3, code_of(x.dtype),
                             stream_ptr(x.device)), "avgpool2")
    return y


def upsample2(x, mode="trilinear", scale=1.0, prev=None, c0=None, c1=None):
    require_cuda(x, prev, c0, c1)
    acc = _acc(x.dtype)
    c0 = None if c0 is None else c0.to(acc).contiguous()
    c1 = None if c1 is None else c1.to(acc).contiguous()
    m1, m2, m3 = _grid(x)
    y = torch.empty(*x.shape[:-3], 2 * m1, 2 * m2, 2 * m3, dtype=x.dtype, device=x.device)
    check(lib.mlpcg_upsample2(ptr(x), ptr(y), _batch(x), 2 * m1, 2 * m2, 2 * m3,
                              int(mode == "nearest"), float(scale), ptr(prev), ptr(c0), ptr(c1),
                              code_of(x.dtype), stream_ptr(x.device)), "upsample2")
    return y


def upsample2_adjoint(g, mode="trilinear", scale=1.0):
    require_cuda(g)
    f1, f2, f3 = _grid(g)
    y = torch.empty(*g.shape[:-3], f1 // 2, f2 // 2, f3 // 2, dtype=g.dtype, device=g.device)
    check(lib.mlpcg_upsample2_adjoint(ptr(g), ptr(y), _batch(g), f1 // 2, f2 // 2, f3 // 2,
                                      int(mode == "nearest"), float(scale), code_of(g.dtype),
                                      stream_ptr(g.device)), "upsample2_adjoint")
    return y


# ------------------------------------------------------------------ Laplacian / dots

def flags_code(flags):
    """flags: int8 tensor (N1, N2, N3) on the GPU -> uint8 code tensor."""
    require_cuda(flags)
    flags = flags.to(torch.int8).contiguous()
    n1, n2, n3 = _grid(flags)
    code = torch.empty(flags.shape, dtype=torch.uint8, device=flags.device)
    check(lib.mlpcg_flags_code(ptr(flags), ptr(code), n1, n2, n3, stream_ptr(flags.device)),
          "flags_code")
    return code


def spmv_flags(code, x):
    require_cuda(code, x)
    n1, n2, n3 = _grid(code)
    y = torch.empty_like(x)
    check(lib.mlpcg_spmv_flags(ptr(code), ptr(x), ptr(y), n1, n2, n3, stream_ptr(x.device)),
          "spmv_flags")
    return y


def loss_resid(code, x, b):
    """Training loss of one right-hand side on the full grid: r = b - M A M x (x float32 or
    bfloat16, any values off the fluid cells M = code != 0) and the (2,) float64 device
    tensor [r.r, b.b]."""
    require_cuda(code, x, b)
    n1, n2, n3 = _grid(code)
    x = x.contiguous()
    b = b.float().contiguous()
    r = torch.empty(n1 * n2 * n3, dtype=torch.float32, device=x.device)
    out = torch.empty(2, dtype=torch.float64, device=x.device)
    check(lib.mlpcg_loss_resid(ptr(x), code_of(x.dtype), ptr(b), ptr(code), ptr(r), n1, n2, n3, ptr(out),
                               ptr(_dot_workspace(x.device)), stream_ptr(x.device)), "loss_resid")
    return r, out


def loss_grad(code, r, coef, dtype):
    """g = ((float)coef * A M r) * M on the full grid, as `dtype` (float32 or bfloat16);
    coef: 0-dim float64 device tensor."""
    require_cuda(code, r, coef)
    n1, n2, n3 = _grid(code)
    coef = coef.reshape(1).contiguous()
    g = torch.empty(n1 * n2 * n3, dtype=dtype, device=r.device)
    check(lib.mlpcg_loss_grad(ptr(r.contiguous()), ptr(code), ptr(coef), ptr(g), code_of(dtype), n1, n2, n3,
                              stream_ptr(r.device)), "loss_grad")
    return g


def comb(c0, x0, c1, x1):
    """c0 * x0 + c1 * x1 for float32 / bfloat16 x0, x1 of one shape and one-element
    coefficient tensors (float32 or bfloat16), with torch's roundings of that expression."""
    require_cuda(c0, x0, c1, x1)
    x0, x1 = x0.contiguous(), x1.contiguous()
    if x0.shape != x1.shape or x0.dtype != x1.dtype:
        raise ValueError("comb: x0 and x1 must have one shape and dtype")
    c0, c1 = _coefs(c0, c1)
    y = torch.empty_like(x0)
    check(lib.mlpcg_comb(ptr(x0), ptr(x1), ptr(c0), ptr(c1), code_of(c0.dtype), ptr(y), x0.numel(),
                         code_of(x0.dtype), stream_ptr(x0.device)), "comb")
    return y


def comb_backward(g, c0, x0, c1, x1, need_g0=True, need_g1=True, out_dtype=torch.float64):
    """Backward of comb: (c0 * g or None, c1 * g or None, (2,) [g.x0, g.x1] accumulated in
    float64 and returned as out_dtype (float64 or float32))."""
    require_cuda(g, c0, x0, c1, x1)
    g, x0, x1 = g.contiguous(), x0.contiguous(), x1.contiguous()
    if not (g.shape == x0.shape == x1.shape and g.dtype == x0.dtype == x1.dtype):
        raise ValueError("comb_backward: g, x0 and x1 must have one shape and dtype")
    c0, c1 = _coefs(c0, c1)
    g0 = torch.empty_like(g) if need_g0 else None
    g1 = torch.empty_like(g) if need_g1 else None
    out = torch.empty(2, dtype=out_dtype, device=g.device)
    check(lib.mlpcg_comb_backward(ptr(g), ptr(x0), ptr(x1), ptr(c0), ptr(c1), code_of(c0.dtype),
                                  ptr(g0) if need_g0 else None, ptr(g1) if need_g1 else None, g.numel(),
                                  code_of(g.dtype), ptr(out), code_of(out_dtype), ptr(_dot_workspace(g.device)),
                                  stream_ptr(g.device)), "comb_backward")
    return g0, g1, out


def _coefs(c0, c1):
    if c0.numel() != 1 or c1.numel() != 1:
        raise ValueError("comb: coefficients must have one element")
    if c0.dtype != c1.dtype:
        c1 = c1.to(c0.dtype)
    return c0.contiguous(), c1.contiguous()


def spmv_csr(rowptr, col, val, x):
    require_cuda(rowptr, col, val, x)
    y = torch.empty_like(x)
    check(lib.mlpcg_spmv_csr(ptr(rowptr), ptr(col), ptr(val), ptr(x), ptr(y), x.numel(),
                             code_of(val.dtype), stream_ptr(x.device)), "spmv_csr")
    return y


_dot_ws = {}


def _dot_workspace(dev):
    ws = _dot_ws.get(dev)
    if ws is None:
        ws = torch.zeros(lib.mlpcg_dot_workspace_size(), dtype=torch.uint8, device=dev)
        _dot_ws[dev] = ws
    return ws


def dot(x, y):
    """FP64-accumulated dot of two float32 vectors -> 0-dim float64 tensor (device)."""
    require_cuda(x, y)
    dev = x.device
    ws = _dot_workspace(dev)
    out = torch.empty((), dtype=torch.float64, device=dev)
    check(lib.mlpcg_dot(ptr(x), ptr(y), x.numel(), ptr(out), ptr(ws), stream_ptr(dev)), "dot")
    return out
