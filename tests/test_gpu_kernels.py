"""GPU parity of the individual kernels (through the C ABI) against the FP64 oracle.

Tolerances: FP64 kernels 1e-11 relative to the output scale (same sums, different
order); FP32 1e-5 relative (the stated preconditioner bar); BF16/FP16 storage only
for the stencil weights, at their rounding (2^-8 / 2^-11 relative)."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import model as M

pytestmark = pytest.mark.gpu


def _rand(shape, seed, dtype=torch.float64):
    g = torch.Generator().manual_seed(seed)
    return (torch.rand(*shape, generator=g, dtype=torch.float64) * 2 - 1).to(dtype)


def _rel(a, b):
    a = a.detach().double().cpu()
    b = b.detach().double().cpu()
    return float((a - b).abs().max() / max(b.abs().max(), 1e-300))


SHAPES = [(8, 8, 8), (6, 10, 12), (4, 6, 5), (5, 3, 16)]


@pytest.mark.parametrize("C", [1, 2, 3])
@pytest.mark.parametrize("shape", SHAPES)
def test_stencil_weights_f64(dev, C, shape):
    from mlpcg_amd import kernels as K
    img, W, b = _rand((C, *shape), 1), _rand((27, C, 3, 3, 3), 2), _rand((27,), 3)
    Kg = K.stencil_weights(img.to(dev), W.to(dev), b.to(dev))
    Kr = M.stencil_weights(img, W, b)
    assert _rel(Kg, Kr) < 1e-12


@pytest.mark.parametrize("C", [1, 3])
@pytest.mark.parametrize("shape", SHAPES)
def test_stencil_weights_transposed_f64(dev, C, shape):
    from mlpcg_amd import kernels as K
    img, W, b = _rand((C, *shape), 4), _rand((27, C, 3, 3, 3), 5), _rand((27,), 6)
    Kt = K.stencil_weights(img.to(dev), W.to(dev), b.to(dev), transposed=True).cpu()
    Kr = M.stencil_weights(img, W, b)
    ref = torch.zeros_like(Kr)
    for q, d in enumerate(M.TAPS):
        ref[q] = M._shifted(Kr[26 - q], d)
    assert _rel(Kt, ref) < 1e-12


@pytest.mark.parametrize("dtype,rnd", [(torch.float32, 0.0), (torch.bfloat16, 2 ** -8),
                                       (torch.float16, 2 ** -11)])
def test_stencil_weights_storage_types(dev, dtype, rnd):
    """FP32 accumulation (error <= 1e-6 of the term scale sum |W||img| + |b|), then one
    rounding to the storage type (<= half an ulp of the value)."""
    from mlpcg_amd import kernels as K
    C, shape = 1, (16, 8, 32)
    img, W, b = _rand((C, *shape), 7).abs(), _rand((27, C, 3, 3, 3), 8), _rand((27,), 9)
    Kg = K.stencil_weights(img.float().to(dev), W.float().to(dev), b.float().to(dev), dtype=dtype)
    assert Kg.dtype == dtype
    Kr = M.stencil_weights(img.float().double(), W.float().double(), b.float().double())
    scale = M.stencil_weights(img.float().double(), W.float().double().abs(), b.float().double().abs())
    err = (Kg.double().cpu() - Kr).abs()
    assert bool((err <= 1e-6 * scale + rnd * Kr.abs()).all())


@pytest.mark.parametrize("trans", [False, True])
@pytest.mark.parametrize("C,B,shape", [(1, 1, (8, 8, 8)), (3, 3, (6, 10, 12)), (2, 5, (4, 6, 5)),
                                       (1, 2, (5, 3, 16))])
def test_sm_block_forward_f64(dev, trans, C, B, shape):
    from mlpcg_amd import kernels as K
    img, x = _rand((C, *shape), 10), _rand((B, 1, *shape), 11)
    W, b = _rand((27, C, 3, 3, 3), 12), _rand((27,), 13)
    f = K.sm_block_trans_forward if trans else K.sm_block_forward
    y = f(img.to(dev), x.to(dev), W.to(dev), b.to(dev))
    yr = (M.sm_block_trans if trans else M.sm_block)(img, x, W, b)
    assert _rel(y, yr) < 1e-12
    # inference entry point: same result
    fi = K.sm_block_trans_inference if trans else K.sm_block_inference
    assert torch.equal(fi(img.to(dev), x.to(dev), W.to(dev), b.to(dev)), y)


@pytest.mark.parametrize("trans", [False, True])
def test_sm_block_forward_f32(dev, trans):
    from mlpcg_amd import kernels as K
    C, B, shape = 3, 2, (16, 16, 16)
    img = (_rand((C, *shape), 14) > 0).float()
    x = _rand((B, 1, *shape), 15).float()
    W, b = _rand((27, C, 3, 3, 3), 16).float() / 9, _rand((27,), 17).float() / 9
    f = K.sm_block_trans_forward if trans else K.sm_block_forward
    y = f(img.to(dev), x.to(dev), W.to(dev), b.to(dev))
    yr = (M.sm_block_trans if trans else M.sm_block)(img.double(), x.double(), W.double(), b.double())
    assert _rel(y, yr) < 1e-5


@pytest.mark.parametrize("trans", [False, True])
@pytest.mark.parametrize("C,B,shape", [(1, 2, (6, 5, 8)), (3, 1, (4, 6, 4)), (2, 3, (5, 4, 6)),
                                       (1, 1, (5, 4, 16)), (3, 1, (4, 3, 8))])
def test_sm_block_backward_f64(dev, trans, C, B, shape):
    """Gradients vs torch autograd of the FP64 oracle."""
    from mlpcg_amd import kernels as K
    img, x, g = _rand((C, *shape), 18), _rand((B, 1, *shape), 19), _rand((B, 1, *shape), 20)
    W, b = _rand((27, C, 3, 3, 3), 21), _rand((27,), 22)
    f = K.sm_block_trans_backward if trans else K.sm_block_backward
    gx, gW, gb = f(g.to(dev), img.to(dev), x.to(dev), W.to(dev), b.to(dev))
    xr, Wr, br = x.clone().requires_grad_(), W.clone().requires_grad_(), b.clone().requires_grad_()
    y = (M.sm_block_trans if trans else M.sm_block)(img, xr, Wr, br)
    (y * g).sum().backward()
    assert _rel(gx, xr.grad) < 1e-11
    assert _rel(gW, Wr.grad) < 1e-11
    assert _rel(gb, br.grad) < 1e-11


@pytest.mark.parametrize("trans", [False, True])
def test_sm_block_backward_finite_differences_f64(dev, trans):
    """Central differences of the GPU forward in FP64 on a 16^3 grid (to 1e-6)."""
    from mlpcg_amd import kernels as K
    C, B, shape = 1, 2, (16, 16, 16)
    img = _rand((C, *shape), 23).abs().to(dev)
    x, R = _rand((B, 1, *shape), 24).to(dev), _rand((B, 1, *shape), 25).to(dev)
    W, b = (_rand((27, C, 3, 3, 3), 26) / 9).to(dev), (_rand((27,), 27) / 9).to(dev)
    fwd = K.sm_block_trans_forward if trans else K.sm_block_forward
    bwd = K.sm_block_trans_backward if trans else K.sm_block_backward
    gx, gW, gb = bwd(R, img, x, W, b)
    L = lambda x_, W_, b_: float((fwd(img, x_, W_, b_) * R).sum())
    h = 1e-4
    for seed in range(3):
        dW, db, dx = _rand(W.shape, 100 + seed).to(dev), _rand(b.shape, 200 + seed).to(dev), \
            _rand(x.shape, 300 + seed).to(dev)
        fd = (L(x, W + h * dW, b) - L(x, W - h * dW, b)) / (2 * h)
        an = float((gW * dW).sum())
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an))
        fd = (L(x, W, b + h * db) - L(x, W, b - h * db)) / (2 * h)
        assert abs(fd - float((gb * db).sum())) <= 1e-6 * max(1.0, abs(fd))
        fd = (L(x + h * dx, W, b) - L(x - h * dx, W, b)) / (2 * h)
        assert abs(fd - float((gx * dx).sum())) <= 1e-6 * max(1.0, abs(fd))


@pytest.mark.parametrize("trans", [False, True])
@pytest.mark.parametrize("B,shape", [(2, (6, 5, 8)), (1, (4, 6, 12)), (3, (3, 4, 20)), (1, (8, 8, 96)),
                                     (2, (5, 4, 6)), (1, (6, 5, 32)), (1, (3, 4, 256))])
def test_sm_block_backward_f32_vs_f64(dev, trans, B, shape):
    """One-channel FP32 backward (matrix-core weight gradient where n3 % 4 == 0, VALU
    otherwise) against the FP64 backward on the same inputs: dx, dW, db."""
    from mlpcg_amd import kernels as K
    img = _rand((1, *shape), 40).abs()
    x, g = _rand((B, 1, *shape), 41), _rand((B, 1, *shape), 42)
    W, b = _rand((27, 1, 3, 3, 3), 43) / 9, _rand((27,), 44) / 9
    f = K.sm_block_trans_backward if trans else K.sm_block_backward
    ref = f(g.to(dev), img.to(dev), x.to(dev), W.to(dev), b.to(dev))
    out = f(g.float().to(dev), img.float().to(dev), x.float().to(dev), W.float().to(dev), b.float().to(dev))
    for o, r in zip(out, ref):
        assert o.dtype == torch.float32
        assert _rel(o, r) < 2e-6


@pytest.mark.parametrize("trans", [False, True])
@pytest.mark.parametrize("C,shape", [(1, (8, 8, 64)), (1, (5, 7, 128)), (2, (4, 6, 64)), (3, (6, 5, 128)),
                                     (3, (1, 3, 64)), (1, (64, 100, 256)), (3, (16, 16, 64))])
def test_sm_block_backward_bf16_weight_grad(dev, trans, C, shape):
    """BF16 weight gradient on the BF16 matrix cores (n3 % 64 == 0): with BF16 g, x and
    image every product g x img is exact and the sums are FP32, so dW and db match the
    FP64 backward of the same (BF16-valued) inputs to FP32 summation error.  The shapes
    cover one-plane grids, ragged row columns (100 rows in columns of 7) and C = 1..3."""
    from mlpcg_amd import kernels as K
    img = _rand((C, *shape), 50).abs().to(torch.bfloat16)
    x, g = _rand((1, 1, *shape), 51).to(torch.bfloat16), _rand((1, 1, *shape), 52).to(torch.bfloat16)
    W, b = _rand((27, C, 3, 3, 3), 53) / 9, _rand((27,), 54) / 9
    f = K.sm_block_trans_backward if trans else K.sm_block_backward
    _, gW, gb = f(g.to(dev), img.to(dev), x.to(dev), W.float().to(dev), b.float().to(dev))
    _, rW, rb = f(g.double().to(dev), img.double().to(dev), x.double().to(dev), W.to(dev), b.to(dev))
    assert gW.dtype == torch.float32
    assert _rel(gW, rW) < 1e-5
    assert _rel(gb, rb) < 1e-5


@pytest.mark.parametrize("trans", [False, True])
@pytest.mark.parametrize("C,shape", [(1, (8, 8, 64)), (3, (5, 7, 128)), (1, (4, 6, 256)), (2, (2, 3, 512)),
                                     (1, (3, 5, 8)), (1, (64, 64, 64)), (2, (64, 32, 128)), (3, (128, 128, 256)),
                                     (1, (256, 128, 256))])
def test_sm_block_backward_bf16_dx_exact(dev, trans, C, shape):
    """BF16 input gradient (the transposed gather, with z-rows of 1..64 runs per wave, and
    the forward apply for the transpose block; on grids of 2^18 voxels or more the stencils
    come from the matrix cores, with 2-, 4- and 8-plane tiles): with exactly representable
    weights, image and output gradient every stencil and every sum is exact, so dx equals
    the FP64 backward rounded to BF16 bit for bit -- except on the block's matrix-core
    grids, whose column sums (1 to 4 tiles of 64 along z) are stored in BF16: there dx is
    within the bound of that rounding."""
    from mlpcg_amd import kernels as K
    g = torch.Generator().manual_seed(31 + C)
    span = 2 if C == 1 else 1
    W = torch.randint(-span, span + 1, (27, C, 3, 3, 3), generator=g).double() / 4
    b = torch.randint(-2, 3, (27,), generator=g).double() / 4
    img = torch.randint(0, 4, (C, *shape), generator=g).double() / 4
    x = torch.randint(-4, 5, (1, 1, *shape), generator=g).double()
    go = torch.randint(-4, 5, (1, 1, *shape), generator=g).double()
    f = K.sm_block_trans_backward if trans else K.sm_block_backward
    dx64, _, _ = f(go.to(dev), img.to(dev), x.to(dev), W.to(dev), b.to(dev))
    dx16, _, _ = f(go.to(dev).bfloat16(), img.to(dev).bfloat16(), x.to(dev).bfloat16(), W.float().to(dev),
                   b.float().to(dev))
    assert dx16.dtype == torch.bfloat16
    n1, n2, n3 = shape
    if not trans and n1 * n2 * n3 >= 1 << 18 and n1 % 8 == 0 and n2 % 8 == 0 and n3 % 64 == 0:
        # the block's matrix-core grids sum g K along z-rows into nine column sums stored in
        # BF16 (one rounding, <= 2^-9 of each sum), then round dx: within 2^-8 of the
        # absolute transposed apply (|g|, |W|, |b| with the non-negative image)
        dxa, _, _ = f(go.abs().to(dev), img.to(dev), x.to(dev), W.abs().to(dev), b.abs().to(dev))
        err = (dx16.double() - dx64).abs()
        assert bool((err <= 2 ** -8 * dxa).all()), float((err - 2 ** -8 * dxa).max())
    else:
        assert torch.equal(dx16, dx64.bfloat16())


@pytest.mark.parametrize("C,shape", [(1, (64, 64, 64)), (2, (64, 32, 128)), (1, (8, 8, 64))])
def test_sm_block_bf16_with_registered_f16_image(dev, C, shape):
    """The BF16 block entry points with the caller's FP16 image copy and packed weights
    (_ex) give the same bits as the plain ones, which convert the image and pack the
    weights themselves (and grids without the matrix-core path ignore both)."""
    from mlpcg_amd import kernels as K
    img = _rand((C, *shape), 70).abs().bfloat16().to(dev)
    img_reg = img.clone()
    K.register_f16_image(img_reg)
    assert K._f16_copy(img_reg, img_reg) is not None and K._f16_copy(img, img) is None
    x = _rand((1, 1, *shape), 71).bfloat16().to(dev)
    go = _rand((1, 1, *shape), 72).bfloat16().to(dev)
    W, b = (_rand((27, C, 3, 3, 3), 73) / 9).float().to(dev), (_rand((27,), 74) / 9).float().to(dev)
    assert torch.equal(K.sm_block_forward(img_reg, x, W, b), K.sm_block_forward(img, x, W, b))
    for u, v in zip(K.sm_block_backward(go, img_reg, x, W, b), K.sm_block_backward(go, img, x, W, b)):
        assert torch.equal(u, v)
    # and with the weights packed once for a stretch of calls (a view sharing the storage
    # finds the packing; none is kept after the block)
    y_ref, bw_ref = K.sm_block_forward(img, x, W, b), K.sm_block_backward(go, img, x, W, b)
    with K.packed_weights([(W, b)]):
        assert K._ex_args(img, x, W.view_as(W), b)[1] is not None
        assert torch.equal(K.sm_block_forward(img_reg, x, W, b), y_ref)
        for u, v in zip(K.sm_block_backward(go, img, x, W, b), bw_ref):
            assert torch.equal(u, v)
    assert K._ex_args(img, x, W, b) is None


@pytest.mark.parametrize("C,shape", [(1, (64, 64, 128)), (3, (64, 32, 192))])
def test_sm_block_backward_bf16_dx_column_sums(dev, C, shape):
    """The block's BF16 input gradient on matrix-core grids (column sums of g K stored in
    BF16, tile ends joined through the side array) with random data: within the rounding
    bound of the absolute transposed apply; the weight and bias gradients are unchanged by
    that path (FP32 against the FP64 backward of the same BF16 inputs)."""
    from mlpcg_amd import kernels as K
    img = _rand((C, *shape), 60).abs().bfloat16().double()
    x = _rand((1, 1, *shape), 61).bfloat16().double()
    go = _rand((1, 1, *shape), 62).bfloat16().double()
    W, b = (_rand((27, C, 3, 3, 3), 63) / 9).float().double(), (_rand((27,), 64) / 9).float().double()
    dx64, gw64, gb64 = K.sm_block_backward(go.to(dev), img.to(dev), x.to(dev), W.to(dev), b.to(dev))
    dxa, _, _ = K.sm_block_backward(go.abs().to(dev), img.to(dev), x.to(dev), W.abs().to(dev), b.abs().to(dev))
    dx16, gw, gb = K.sm_block_backward(go.bfloat16().to(dev), img.bfloat16().to(dev), x.bfloat16().to(dev),
                                       W.float().to(dev), b.float().to(dev))
    err = (dx16.double() - dx64).abs()
    # FP16 image copy in the stencil product (exact for these BF16 values in [0, 1]) and
    # FP32 sums: the BF16 roundings of the column sums and of dx dominate
    assert bool((err <= 2 ** -8 * dxa + 1e-6 * float(dxa.max())).all())
    assert _rel(gw, gw64) < 1e-5 and _rel(gb, gb64) < 1e-5


@pytest.mark.parametrize("C,shape", [(1, (8, 8, 8)), (3, (6, 10, 12)), (2, (1, 4, 5)), (1, (16, 16, 32))])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_sm_linear(dev, C, shape, dtype):
    from mlpcg_amd import kernels as K
    img, w, b = _rand((C, *shape), 28).to(dtype), _rand((1, C, 3, 3, 3), 29).to(dtype), _rand((1,), 30).to(dtype)
    z, ybar = K.sm_linear(img.to(dev), w.to(dev), b.to(dev))
    zr, ybr = M.sm_linear(img.double(), w.double(), b.double())
    tol = 1e-12 if dtype == torch.float64 else 1e-6
    assert _rel(ybar, ybr) < tol
    assert abs(float(z) - float(zr)) < tol * max(1.0, float(zr.abs()))
    g = torch.tensor([1.7], dtype=z.dtype, device=dev)
    gw, gb = K.sm_linear_backward(g, ybar)
    assert torch.allclose(gw, 1.7 * ybar) and float(gb) == pytest.approx(1.7)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32, torch.bfloat16, torch.float16])
@pytest.mark.parametrize("shape", [(8, 8, 8), (4, 6, 10), (2, 2, 2)])
def test_avgpool2(dev, dtype, shape):
    from mlpcg_amd import kernels as K
    x = _rand((3, 1, *shape), 31).to(dtype)
    y = K.avgpool2(x.to(dev))
    yr = F.avg_pool3d(x.double(), (2, 2, 2))
    tol = {torch.float64: 1e-15, torch.float32: 5e-7, torch.bfloat16: 2 ** -8, torch.float16: 2 ** -11}[dtype]
    assert _rel(y, yr) < tol


@pytest.mark.parametrize("mode", ["trilinear", "nearest"])
@pytest.mark.parametrize("shape", [(4, 4, 4), (2, 3, 5), (1, 1, 1), (8, 4, 2)])
def test_upsample2_f64(dev, mode, shape):
    from mlpcg_amd import kernels as K
    x = _rand((2, 1, *shape), 32)
    y = K.upsample2(x.to(dev), mode, 1.0)
    yr = F.interpolate(x, scale_factor=2, mode=mode) if mode != "nearest" else F.interpolate(x, scale_factor=2)
    assert _rel(y, yr) < 1e-14
    # fused axpby form and scale
    prev = _rand(y.shape, 33).to(dev)
    c0 = torch.tensor([0.3], dtype=torch.float64, device=dev)
    c1 = torch.tensor([-1.5], dtype=torch.float64, device=dev)
    y2 = K.upsample2(x.to(dev), mode, 0.125, prev=prev, c0=c0, c1=c1)
    assert _rel(y2, 0.3 * prev - 1.5 * 0.125 * y) < 1e-14


@pytest.mark.parametrize("mode", ["trilinear", "nearest"])
@pytest.mark.parametrize("shape", [(4, 4, 4), (2, 3, 5), (1, 2, 1)])
def test_upsample2_adjoint_f64(dev, mode, shape):
    """<U x, y> = <x, U^T y>."""
    from mlpcg_amd import kernels as K
    x = _rand((2, 1, *shape), 34).to(dev)
    yf = _rand((2, 1, *(2 * s for s in shape)), 35).to(dev)
    lhs = float((K.upsample2(x, mode, 0.5) * yf).sum())
    rhs = float((x * K.upsample2_adjoint(yf, mode, 0.5)).sum())
    assert abs(lhs - rhs) < 1e-12 * max(1.0, abs(lhs))


@pytest.mark.parametrize("mode", ["trilinear", "nearest"])
@pytest.mark.parametrize("dtype,tol", [(torch.float32, 1e-6), (torch.bfloat16, 2 ** -8),
                                       (torch.float16, 2 ** -11)])
@pytest.mark.parametrize("shape", [(128, 128, 64), (64, 64, 130)])
def test_upsample2_adjoint_runs(dev, mode, dtype, tol, shape):
    """Coarse grids of >= 2^19 cells take the z-run gather (4 or 2 coarse cells per thread);
    per element it equals the transpose of torch's upsampling in FP64 up to the output's
    rounding."""
    from mlpcg_amd import kernels as K
    yf = _rand((1, 1, *(2 * s for s in shape)), 36).to(dtype)
    xr = torch.zeros((1, 1, *shape), dtype=torch.float64, requires_grad=True)
    up = F.interpolate(xr, scale_factor=2, mode=mode) if mode != "nearest" else F.interpolate(xr, scale_factor=2)
    up.backward(yf.double())
    ref = xr.grad * 0.5
    out = K.upsample2_adjoint(yf.to(dev), mode, 0.5).double().cpu()
    assert out.shape == ref.shape
    err = (out - ref).abs()
    # output rounding per element, FP32 accumulation over at most 64 products
    assert float((err - tol * ref.abs()).max()) <= 1e-5 * float(ref.abs().max())


@pytest.mark.parametrize("dtype,tol", [(torch.float32, 1e-6), (torch.bfloat16, 2 ** -7)])
def test_sm_apply_epilogue(dev, dtype, tol):
    """y = c0 * prev + c1 * (scale * stencil), masked before and after."""
    from mlpcg_amd import kernels as K
    shape = (8, 8, 16)
    img, W, b = _rand((1, *shape), 36).abs(), _rand((27, 1, 3, 3, 3), 37) / 9, _rand((27,), 38) / 9
    x = _rand((1, 1, *shape), 39)
    prev = _rand((1, 1, *shape), 40)
    ma = (_rand(shape, 41) > -0.5).to(torch.uint8)
    mb = (_rand(shape, 42) > -0.5).to(torch.uint8)
    Kd = K.stencil_weights(img.float().to(dev), W.float().to(dev), b.float().to(dev), dtype=dtype)
    scale = torch.tensor(0.25, dtype=torch.float64, device=dev)
    c0 = torch.tensor([0.7], device=dev)
    c1 = torch.tensor([-1.3], device=dev)
    y = K.sm_apply(Kd, x.float().to(dev), out_dtype=dtype, prev=prev.to(dtype).to(dev), c0=c0, c1=c1,
                   mask_a=ma.to(dev), mask_b=mb.to(dev), in_scale=scale)
    ref = 0.25 * M.sm_block(img.float().double(), x.float().double(), W.float().double(), b.float().double())
    ref = torch.where(ma.bool(), ref, torch.zeros(()))
    ref = 0.7 * prev.to(dtype).double() - 1.3 * ref
    ref = torch.where(mb.bool(), ref, torch.zeros(()))
    assert _rel(y, ref) < tol


def test_dot_f64_accumulation(dev):
    from mlpcg_amd import kernels as K
    n = 1_000_003
    x = _rand((n,), 43).float()
    y = _rand((n,), 44).float()
    d = float(K.dot(x.to(dev), y.to(dev)))
    ref = float((x.double() * y.double()).sum())
    assert abs(d - ref) < 1e-12 * np.sqrt(n)
