"""Drop-in replacement of the reference's sm_model_3d module (sm_model_3d.py:1-444).

Same class names, constructor signatures, parameter names/shapes (so reference
state_dicts load) and forward / eval_forward semantics; the compute goes to the gfx950
kernels of libmlpcg_hip.so through the C ABI.  Differences from the reference, all
internal: inputs are not pre-padded (the kernels handle the halo); the stencil weights
K are recomputed in backward instead of being kept from forward; the weight gradient
is a deterministic block reduction instead of global atomics; any batch size works.

The extension namespaces `smblock3d`, `smblocktrans3d`, `smlinear3d` mirror the
reference's pybind modules (sm_block_3d.cpp:72-76, sm_block_trans_3d.cpp:71-75,
sm_linear_3d.cpp:56-60) on UNPADDED tensors.
"""
from __future__ import annotations

import types

import torch
import torch.nn as nn
import torch.nn.functional as F

from . import kernels as K
from .model import BaseModel

# ---------------------------------------------------------------------------
# Extension namespaces (reference pybind modules, unpadded inputs)
# ---------------------------------------------------------------------------
smblock3d = types.SimpleNamespace(
    inference=lambda image, x, w, b: [K.sm_block_inference(image, x, w, b)],
    forward=lambda image, x, w, b: [K.sm_block_forward(image, x, w, b)],
    backward=lambda g, image, x, w, b: list(K.sm_block_backward(g, image, x, w, b)),
)
smblocktrans3d = types.SimpleNamespace(
    inference=lambda image, x, w, b: [K.sm_block_trans_inference(image, x, w, b)],
    forward=lambda image, x, w, b: [K.sm_block_trans_forward(image, x, w, b)],
    backward=lambda g, image, x, w, b: list(K.sm_block_trans_backward(g, image, x, w, b)),
)
smlinear3d = types.SimpleNamespace(
    inference=lambda image, w, b: list(K.sm_linear(image, w, b)),
    forward=lambda image, w, b: list(K.sm_linear(image, w, b)),
    backward=lambda g, ybar: list(K.sm_linear_backward(g, ybar)),
)


def _tstart(timer, name):
    if timer:
        timer.start(name)


def _tstop(timer, name):
    if timer:
        torch.cuda.synchronize()
        timer.stop(name)


######################
# SM block
######################
class SMBlockFunction3D(torch.autograd.Function):
    """sm_model_3d.py:21-48."""

    @staticmethod
    def inference(image, x, weights, bias, timer=None):
        _tstart(timer, "Forward")
        y, = smblock3d.inference(image, x, weights, bias)
        _tstop(timer, "Forward")
        return y

    @staticmethod
    def forward(ctx, image, x, weights, bias):
        y, = smblock3d.forward(image, x, weights, bias)
        ctx.save_for_backward(image, x, weights, bias)
        return y

    @staticmethod
    def backward(ctx, grad_output):
        image, x, weights, bias = ctx.saved_tensors
        grad_x, grad_w, grad_b = smblock3d.backward(grad_output.contiguous(), image, x, weights, bias)
        return None, grad_x, grad_w, grad_b


class SMBlockTransFunction3D(torch.autograd.Function):
    """sm_model_3d.py:50-76."""

    @staticmethod
    def inference(image, x, weights, bias, timer=None):
        _tstart(timer, "Forward")
        y, = smblocktrans3d.inference(image, x, weights, bias)
        _tstop(timer, "Forward")
        return y

    @staticmethod
    def forward(ctx, image, x, weights, bias):
        y, = smblocktrans3d.forward(image, x, weights, bias)
        ctx.save_for_backward(image, x, weights, bias)
        return y

    @staticmethod
    def backward(ctx, grad_output):
        image, x, weights, bias = ctx.saved_tensors
        grad_x, grad_w, grad_b = smblocktrans3d.backward(grad_output.contiguous(), image, x, weights,
                                                         bias)
        return None, grad_x, grad_w, grad_b


class SmallSMBlock3D(BaseModel):
    """sm_model_3d.py:78-94."""

    def __init__(self, num_imgs=3, mask=False):
        super().__init__()
        self.mask = mask
        self.weight = nn.Parameter(torch.ones(27, num_imgs, 3, 3, 3))
        self.bias = nn.Parameter(torch.ones(27))
        self.reset_parameters(self.weight, self.bias)

    def forward(self, image, x):
        y = SMBlockFunction3D.apply(image, x, self.weight, self.bias)
        if self.mask:
            y = torch.where(image[1] == 0, torch.zeros((), dtype=y.dtype, device=y.device), y)
        return y

    def eval_forward(self, image, x, timer=None):
        y = SMBlockFunction3D.inference(image, x, self.weight, self.bias, timer)
        if self.mask:
            y = torch.where(image[1] == 0, torch.zeros((), dtype=y.dtype, device=y.device), y)
        return y


class SmallSMBlockTrans3D(BaseModel):
    """sm_model_3d.py:96-112 (weight tying through the weight/bias arguments)."""

    def __init__(self, num_imgs=3, weight=None, bias=None):
        super().__init__()
        self.weight = nn.Parameter(torch.ones(27, num_imgs, 3, 3, 3)) if weight is None else weight
        self.bias = nn.Parameter(torch.ones(27)) if bias is None else bias
        if weight is None and bias is None:
            self.reset_parameters(self.weight, self.bias)

    def forward(self, image, x):
        return SMBlockTransFunction3D.apply(image, x, self.weight, self.bias)

    def eval_forward(self, image, x, timer=None):
        return SMBlockTransFunction3D.inference(image, x, self.weight, self.bias, timer)


######################
# SM linear
######################
class SMLinearFunction3D(torch.autograd.Function):
    """sm_model_3d.py:118-144."""

    @staticmethod
    def inference(image, weights, bias, timer=None):
        _tstart(timer, "Forward")
        z, _ = smlinear3d.inference(image, weights, bias)
        _tstop(timer, "Forward")
        return z

    @staticmethod
    def forward(ctx, image, weights, bias, timer=None):
        z, ybar = smlinear3d.forward(image, weights, bias)
        ctx.save_for_backward(ybar)
        ctx.wdtype = (weights.dtype, bias.dtype)
        return z

    @staticmethod
    def backward(ctx, grad_output):
        ybar, = ctx.saved_tensors
        grad_w, grad_b = smlinear3d.backward(grad_output, ybar)
        return None, grad_w.to(ctx.wdtype[0]), grad_b.to(ctx.wdtype[1]), None


class SmallLinearBlock3DNew(BaseModel):
    """sm_model_3d.py:146-157: parameters are the mean over 27 of a block-shaped init."""

    def __init__(self, num_imgs=3):
        super().__init__()
        weight = torch.ones(27, num_imgs, 3, 3, 3)
        bias = torch.ones(27)
        self.reset_parameters(weight, bias)
        self.weight = nn.Parameter(weight.mean(dim=0, keepdim=True))
        self.bias = nn.Parameter(bias.mean(dim=0, keepdim=True))

    def forward(self, image):
        return SMLinearFunction3D.apply(image, self.weight, self.bias)

    def eval_forward(self, image, timer=None):
        return SMLinearFunction3D.inference(image, self.weight, self.bias, timer)


######################
# Level transfers (F.avg_pool3d / F.interpolate replacements with HIP backward)
######################
class AvgPool2Function(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        return K.avgpool2(x.contiguous())

    @staticmethod
    def backward(ctx, g):
        return K.upsample2(g.contiguous(), "nearest", 0.125)


class Upsample2Function(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, mode, scale):
        ctx.mode, ctx.scale = mode, scale
        return K.upsample2(x.contiguous(), mode, scale)

    @staticmethod
    def backward(ctx, g):
        return K.upsample2_adjoint(g.contiguous(), ctx.mode, ctx.scale), None, None


def avg_pool(x):
    """F.avg_pool3d(x, (2, 2, 2))."""
    if torch.is_grad_enabled() and x.requires_grad:
        return AvgPool2Function.apply(x)
    return K.avgpool2(x.contiguous())


def upsample(x, mode="trilinear", scale=1.0):
    """scale * F.interpolate(x, scale_factor=2, mode=mode) (align_corners=False)."""
    if mode not in ("trilinear", "nearest"):
        raise ValueError(f"unsupported interpolation mode {mode}")
    if torch.is_grad_enabled() and x.requires_grad:
        return Upsample2Function.apply(x, mode, scale)
    return K.upsample2(x.contiguous(), mode, scale)


class _CombFunction(torch.autograd.Function):
    """c0 * x0 + c1 * x1 with coefficients of shape (1,) (sm_model_3d.py:245,258,282,286).
    FP32/BF16 levels run it as one library kernel each way (forward: one pass with torch's
    roundings of the expression; backward: c0 g, c1 g and the coefficient gradients g.x0,
    g.x1 in FP64 from one pass), instead of three and six torch kernels.  The coefficient
    gradients are deterministic two-kernel reductions, so a training step that contains
    them can be captured in a HIP graph (a torch reduction over many workgroups records
    memset nodes, which give wrong results on replays after the first)."""

    @staticmethod
    def forward(ctx, c0, x0, c1, x1):
        ctx.save_for_backward(c0, x0, c1, x1)
        if _comb_native(c0, x0, c1, x1):
            return K.comb(c0, x0, c1, x1)
        return c0.to(x0.dtype) * x0 + c1.to(x1.dtype) * x1

    @staticmethod
    def backward(ctx, g):
        c0, x0, c1, x1 = ctx.saved_tensors
        need = ctx.needs_input_grad
        if _comb_native(c0, x0, c1, x1) and g.dtype == x0.dtype:
            # the coefficient gradients straight in the coefficients' type when that is FP32
            odt = torch.float32 if c0.dtype == c1.dtype == torch.float32 else torch.float64
            g0, g1, dots = K.comb_backward(g, c0, x0, c1, x1, need[1], need[3], out_dtype=odt)
            return (dots[0:1].reshape(c0.shape).to(c0.dtype) if need[0] else None, g0,
                    dots[1:2].reshape(c1.shape).to(c1.dtype) if need[2] else None, g1)

        def coef_grad(x, c):
            if x.dtype == torch.float64:
                return (g.double() * x).sum().reshape(c.shape).to(c.dtype)
            gf = g.float().contiguous().reshape(-1)
            return K.dot(gf, x.float().contiguous().reshape(-1)).reshape(c.shape).to(c.dtype)

        return (coef_grad(x0, c0) if need[0] else None,
                c0.to(g.dtype) * g if need[1] else None,
                coef_grad(x1, c1) if need[2] else None,
                c1.to(g.dtype) * g if need[3] else None)


def _comb_native(c0, x0, c1, x1):
    """The library's combination kernels cover FP32 / BF16 levels with FP32 / BF16
    coefficients (FP64 stays on torch, where the oracle comparisons run)."""
    return (x0.is_cuda and x0.dtype in (torch.float32, torch.bfloat16) and x1.dtype == x0.dtype
            and x0.shape == x1.shape
            and c0.dtype in (torch.float32, torch.bfloat16) and c1.dtype in (torch.float32, torch.bfloat16)
            and c0.numel() == 1 and c1.numel() == 1)


def _comb(c0, x0, c1, x1):
    return _CombFunction.apply(c0, x0, c1, x1)


def _sm_image(image, x):
    return image if image.dtype == x.dtype else image.to(x.dtype)


######################
# Full SM Model
######################
class SmallSMModelDn3D(BaseModel):
    """sm_model_3d.py:164-287."""

    def __init__(self, n, num_imgs=3, interpolation_mode="trilinear", mask=False, swap_sm_order=False):
        super().__init__()
        self.n = n
        self.mode = interpolation_mode
        self.swap_ord = swap_sm_order
        self.num_imgs = num_imgs
        self.mask = mask
        self.pre = nn.ModuleList([SmallSMBlock3D(num_imgs) for _ in range(n - 1)])
        self.pre.insert(0, SmallSMBlock3D(num_imgs, mask))
        self.post = nn.ModuleList([SmallSMBlock3D(num_imgs) for _ in range(n - 1)])
        self.post.insert(0, SmallSMBlock3D(num_imgs, mask))
        self.l = SmallSMBlock3D(num_imgs)
        self.c0 = nn.ModuleList([SmallLinearBlock3DNew(num_imgs) for _ in range(n)])
        self.c1 = nn.ModuleList([SmallLinearBlock3DNew(num_imgs) for _ in range(n)])

    def eval_forward(self, image, b, timer=None, imgs=[], c0=[], c1=[]):  # noqa: B006 (reference API)
        image = _sm_image(image, b)
        imgs_cached = bool(imgs)
        if not imgs_cached:
            imgs.extend([None] * self.n)
        c0_cached = bool(c0)
        if not c0_cached:
            c0.extend([None] * self.n)
        c1_cached = bool(c1)
        if not c1_cached:
            c1.extend([None] * self.n)
        x = [self.pre[0].eval_forward(image, b, timer)]
        if not imgs_cached:
            imgs[0] = image
        for i in range(1, self.n):
            x.append(avg_pool(x[-1]))
            if not imgs_cached:
                imgs[i] = avg_pool(imgs[i - 1])
            x[-1] = self.pre[i].eval_forward(imgs[i], x[-1], timer)
        x.append(avg_pool(x[-1]))
        fine_img = avg_pool(imgs[-1])
        x[-1] = self.l.eval_forward(fine_img, x[-1], timer)
        for i in range(self.n, 0, -1):
            x[i] = upsample(x[i], self.mode)
            if not c0_cached:
                c0[i - 1] = self.c0[i - 1].eval_forward(imgs[i - 1], timer)
            if not c1_cached:
                c1[i - 1] = self.c1[i - 1].eval_forward(imgs[i - 1], timer)
            if self.swap_ord:
                x[i - 1] = _comb(c0[i - 1], x[i - 1], c1[i - 1], x[i])
                x[i - 1] = self.post[i - 1].eval_forward(imgs[i - 1], x[i - 1], timer)
            else:
                x[i] = self.post[i - 1].eval_forward(imgs[i - 1], x[i], timer)
                x[i - 1] = _comb(c0[i - 1], x[i - 1], c1[i - 1], x[i])
        return x[0]

    def forward(self, image, b):
        image = _sm_image(image, b)
        x = [self.pre[0](image, b)]
        imgs = [image]
        for i in range(1, self.n):
            x.append(avg_pool(x[-1]))
            imgs.append(avg_pool(imgs[-1]))
            x[-1] = self.pre[i](imgs[-1], x[-1])
        x.append(avg_pool(x[-1]))
        imgs.append(avg_pool(imgs[-1]))
        x[-1] = self.l(imgs[-1], x[-1])
        for i in range(self.n, 0, -1):
            x[i] = upsample(x[i], self.mode)
            c0 = self.c0[i - 1](imgs[i - 1])
            c1 = self.c1[i - 1](imgs[i - 1])
            if self.swap_ord:
                x[i - 1] = _comb(c0, x[i - 1], c1, x[i])
                x[i - 1] = self.post[i - 1](imgs[i - 1], x[i - 1])
            else:
                x[i] = self.post[i - 1](imgs[i - 1], x[i])
                x[i - 1] = _comb(c0, x[i - 1], c1, x[i])
        return x[0]

    # -- training with a fixed scene image ------------------------------------------
    def image_cache(self, image, dtype=None):
        """The parts of forward() that depend on the image only: the pyramid (in `dtype`,
        the activation type; forward casts the image to it; BF16 pyramids also keep the
        FP16 copies the matrix-core blocks read) and the SMLinear mean patches ybar of
        levels 0..n-1 (sm_linear_3d_kernel.cu:62-83).  A training loop that reuses a scene
        image computes them once and calls forward_cached."""
        with torch.no_grad():
            imgs = [image.to(dtype or image.dtype).contiguous()]
            for _ in range(self.n):
                imgs.append(K.avgpool2(imgs[-1]))
            if imgs[0].dtype == torch.bfloat16 and imgs[0].is_cuda:
                for im in imgs:  # the FP16 copies the BF16 matrix-core blocks read
                    K.register_f16_image(im)
            ybars = [K.sm_linear(imgs[i], self.c0[i].weight, self.c0[i].bias)[1]
                     for i in range(self.n)]
            # rows of the c0 (levels 0..n-1) then c1 blocks: ybar of their level
            Y = torch.stack([y.reshape(-1) for y in ybars + ybars])
        return imgs, ybars, Y

    def linear_stack(self, dtype):
        """The 2n SMLinear weights (rows) and biases, c0 blocks then c1, in `dtype`: the
        operands of forward_cached's coefficient product.  A step over several scenes builds
        them once and passes them to each scene's forward_cached, so the scenes' gradients
        meet in one tensor before they reach the 4n parameters."""
        lins = list(self.c0) + list(self.c1)
        Wm = torch.stack([m.weight.reshape(-1) for m in lins]).to(dtype)
        bv = torch.cat([m.bias for m in lins]).to(dtype)
        return Wm, bv

    def forward_cached(self, cache, b, lin=None):
        """forward(image, b) with the image pyramid and mean patches of image_cache: same
        result.  All 2n coefficients c = w . ybar + bias come from one batched product
        (its gradients are smlinear3d.backward's: gw = g ybar, gb = g); lin: the operands
        of linear_stack(Y.dtype), if already built."""
        imgs, ybars, Y = cache
        Wm, bv = lin if lin is not None else self.linear_stack(Y.dtype)
        coef = (Wm * Y).sum(1) + bv
        n = self.n
        # one split (its backward is one concatenation of the 2n coefficient gradients,
        # where 2n slices would each scatter into a zero vector that is then summed)
        cs = coef.split(1)

        def lin(which, i):
            return cs[which * n + i]

        x = [self.pre[0](imgs[0], b)]
        for i in range(1, self.n):
            x.append(avg_pool(x[-1]))
            x[-1] = self.pre[i](imgs[i], x[-1])
        x.append(avg_pool(x[-1]))
        x[-1] = self.l(imgs[self.n], x[-1])
        for i in range(self.n, 0, -1):
            x[i] = upsample(x[i], self.mode)
            c0 = lin(0, i - 1)
            c1 = lin(1, i - 1)
            if self.swap_ord:
                x[i - 1] = _comb(c0, x[i - 1], c1, x[i])
                x[i - 1] = self.post[i - 1](imgs[i - 1], x[i - 1])
            else:
                x[i] = self.post[i - 1](imgs[i - 1], x[i])
                x[i - 1] = _comb(c0, x[i - 1], c1, x[i])
        return x[0]

    # -- native plan ----------------------------------------------------------------
    def plan_params(self):
        blocks = list(self.pre) + list(self.post) + [self.l]
        lin = list(self.c0) + list(self.c1)
        return [m.weight for m in blocks + lin], [m.bias for m in blocks + lin]

    def make_preconditioner(self, image, dtype=torch.float32, fused_mode=1, coarse_tail=True, fold_third=False):
        """Native plan with K, the image pyramid and c0/c1 precomputed for `image`;
        fused_mode / coarse_tail / fold_third are the plan's execution policies
        (precond.Preconditioner)."""
        from .precond import Preconditioner
        return Preconditioner(self, image, dtype, fused_mode=fused_mode, coarse_tail=coarse_tail,
                              fold_third=fold_third)


####################################
# SPD SM Model (Symmetric model 1)
####################################
class SPDSMModelDn3D(BaseModel):
    """sm_model_3d.py:293-383 (nearest upsampling / 8, transpose blocks tied to l0/l1/l)."""

    def __init__(self, n, num_imgs=3):
        super().__init__()
        self.n = n
        self.num_imgs = num_imgs
        self.l0 = nn.ModuleList()
        self.l1 = nn.ModuleList()
        self.l0_t = nn.ModuleList()
        self.l1_t = nn.ModuleList()
        for _ in range(n):
            self.l0.append(SmallSMBlock3D(num_imgs))
            self.l1.append(SmallSMBlock3D(num_imgs))
            self.l0_t.append(SmallSMBlockTrans3D(num_imgs, weight=self.l0[-1].weight, bias=self.l0[-1].bias))
            self.l1_t.append(SmallSMBlockTrans3D(num_imgs, weight=self.l1[-1].weight, bias=self.l1[-1].bias))
        self.l = SmallSMBlock3D(num_imgs)
        self.l_t = SmallSMBlockTrans3D(num_imgs, weight=self.l.weight, bias=self.l.bias)
        self.c0 = nn.ModuleList([SmallLinearBlock3DNew(num_imgs) for _ in range(n)])
        self.c1 = nn.ModuleList([SmallLinearBlock3DNew(num_imgs) for _ in range(n)])

    def _run(self, image, b, ev):
        blk = (lambda m, img, x: m.eval_forward(img, x)) if ev else (lambda m, img, x: m(img, x))
        lin = (lambda m, img: m.eval_forward(img)) if ev else (lambda m, img: m(img))
        image = _sm_image(image, b)
        x = [blk(self.l0[0], image, b)]
        imgs = [image]
        c0, c1 = [], []
        for i in range(1, self.n):
            x.append(avg_pool(x[-1]))
            imgs.append(avg_pool(imgs[-1]))
            x[-1] = blk(self.l0[i], imgs[-1], x[-1])
        x.append(avg_pool(x[-1]))
        imgs.append(avg_pool(imgs[-1]))
        x[-1] = blk(self.l, imgs[-1], x[-1])
        for i in range(self.n, 0, -1):
            x[i] = upsample(x[i], "nearest", 0.125)
            c0.insert(0, lin(self.c0[i - 1], imgs[i - 1]))
            c1.insert(0, lin(self.c1[i - 1], imgs[i - 1]))
            x[i] = blk(self.l1[i - 1], imgs[i - 1], x[i])
            x[i - 1] = _comb(c0[0], x[i - 1], c1[0], x[i])
        bb = x[0]
        x = [bb]
        for i in range(self.n):
            x.append(blk(self.l1_t[i], imgs[i], x[i]))
            x[-1] = avg_pool(x[-1])
        x[-1] = blk(self.l_t, imgs[-1], x[-1])
        for i in range(self.n - 1, -1, -1):
            x[i + 1] = upsample(x[i + 1], "nearest", 0.125)
            x[i] = _comb(c0[i], x[i], c1[i], x[i + 1])
            x[i] = blk(self.l0_t[i], imgs[i], x[i])
        return x[0]

    def eval_forward(self, image, b, timer=None, imgs=[], c0=[], c1=[]):  # noqa: B006
        return self._run(image, b, True)

    def forward(self, image, b):
        return self._run(image, b, False)

    def plan_params(self):
        blocks = list(self.l0) + list(self.l1) + [self.l]
        lin = list(self.c0) + list(self.c1)
        return [m.weight for m in blocks + lin], [m.bias for m in blocks + lin]

    def make_preconditioner(self, image, dtype=torch.float32, fused_mode=1, coarse_tail=True):
        from .precond import Preconditioner
        return Preconditioner(self, image, dtype, fused_mode=fused_mode, coarse_tail=coarse_tail)


##########################################
# Reduced SPD model (Symmetric model 2)
##########################################
class SmallSPDSMModelDn3D(BaseModel):
    """sm_model_3d.py:388-444."""

    def __init__(self, n, num_imgs=3):
        super().__init__()
        self.n = n
        self.num_imgs = num_imgs
        self.l0 = nn.ModuleList()
        self.l0_t = nn.ModuleList()
        for _ in range(n):
            self.l0.append(SmallSMBlock3D(num_imgs))
            self.l0_t.append(SmallSMBlockTrans3D(num_imgs, weight=self.l0[-1].weight, bias=self.l0[-1].bias))
        self.l = SmallSMBlock3D(num_imgs)
        self.l_t = SmallSMBlockTrans3D(num_imgs, weight=self.l.weight, bias=self.l.bias)
        self.c0 = nn.ModuleList([SmallLinearBlock3DNew(num_imgs) for _ in range(n)])
        self.c1 = nn.ModuleList([SmallLinearBlock3DNew(num_imgs) for _ in range(n)])

    def _run(self, image, b, ev):
        blk = (lambda m, img, x: m.eval_forward(img, x)) if ev else (lambda m, img, x: m(img, x))
        lin = (lambda m, img: m.eval_forward(img)) if ev else (lambda m, img: m(img))
        image = _sm_image(image, b)
        x = [blk(self.l0[0], image, b)]
        imgs = [image]
        for i in range(1, self.n):
            x.append(avg_pool(x[-1]))
            imgs.append(avg_pool(imgs[-1]))
            x[-1] = blk(self.l0[i], imgs[-1], x[-1])
        x.append(avg_pool(x[-1]))
        imgs.append(avg_pool(imgs[-1]))
        x[-1] = blk(self.l, imgs[-1], x[-1])
        x[-1] = blk(self.l_t, imgs[-1], x[-1])
        for i in range(self.n, 0, -1):
            x[i] = upsample(x[i], "nearest", 0.125)
            c0 = lin(self.c0[i - 1], imgs[i - 1])
            c1 = lin(self.c1[i - 1], imgs[i - 1])
            x[i - 1] = _comb(c0 ** 2, x[i - 1], c1 ** 2, x[i])
            x[i - 1] = blk(self.l0_t[i - 1], imgs[i - 1], x[i - 1])
        return x[0]

    def forward(self, image, b):
        return self._run(image, b, False)

    def eval_forward(self, image, b, timer=None, imgs=[], c0=[], c1=[]):  # noqa: B006
        return self._run(image, b, True)

    def plan_params(self):
        blocks = list(self.l0) + [self.l]
        lin = list(self.c0) + list(self.c1)
        return [m.weight for m in blocks + lin], [m.bias for m in blocks + lin]

    def make_preconditioner(self, image, dtype=torch.float32, fused_mode=1, coarse_tail=True):
        from .precond import Preconditioner
        return Preconditioner(self, image, dtype, fused_mode=fused_mode, coarse_tail=coarse_tail)


__all__ = [
    "smblock3d", "smblocktrans3d", "smlinear3d",
    "SMBlockFunction3D", "SMBlockTransFunction3D", "SMLinearFunction3D",
    "SmallSMBlock3D", "SmallSMBlockTrans3D", "SmallLinearBlock3DNew",
    "SmallSMModelDn3D", "SPDSMModelDn3D", "SmallSPDSMModelDn3D",
    "AvgPool2Function", "Upsample2Function", "avg_pool", "upsample", "F",
]
