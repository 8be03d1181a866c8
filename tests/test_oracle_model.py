"""Pins of the FP64 oracle (oracle/model.py) against brute force, library routines and
closed forms -- chosen so a dropped term, a wrong sign/index or a transposed operand
fails at least one of them."""
import itertools
import math

import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import model as M


@pytest.fixture(autouse=True)
def _float64_default():
    old = torch.get_default_dtype()
    torch.set_default_dtype(torch.float64)
    yield
    torch.set_default_dtype(old)


def _rand(shape, seed):
    g = torch.Generator().manual_seed(seed)
    return torch.rand(*shape, generator=g, dtype=torch.float64) * 2 - 1


@pytest.mark.parametrize("C", [1, 3])
@pytest.mark.parametrize("shape", [(3, 4, 5), (4, 4, 4)])
def test_sm_block_vs_brute_force(C, shape):
    img = _rand((C, *shape), 1)
    x = _rand((2, 1, *shape), 2)
    W = _rand((27, C, 3, 3, 3), 3)
    b = _rand((27,), 4)
    y = M.sm_block(img, x, W, b)
    yb = M.brute_sm_block(img, x, W, b)
    assert torch.allclose(y, yb, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("C", [1, 3])
def test_sm_block_trans_vs_brute_force(C):
    shape = (3, 4, 5)
    img = _rand((C, *shape), 5)
    x = _rand((2, 1, *shape), 6)
    W = _rand((27, C, 3, 3, 3), 7)
    b = _rand((27,), 8)
    y = M.sm_block_trans(img, x, W, b)
    yb = M.brute_sm_block_trans(img, x, W, b)
    assert torch.allclose(y, yb, rtol=1e-12, atol=1e-12)


def test_sm_block_unfold_form():
    """Same sum as the Conv3d + unfold form of sm_model_3d_py.py:15-22, written here with
    torch library routines (F.conv3d, Tensor.unfold)."""
    C, shape = 3, (5, 6, 7)
    img = _rand((C, *shape), 9)
    x = _rand((3, 1, *shape), 10)
    W = _rand((27, C, 3, 3, 3), 11)
    b = _rand((27,), 12)
    K = F.conv3d(img[None], W, b, padding=1)[0].permute(1, 2, 3, 0).unflatten(3, (3, 3, 3))
    xu = F.pad(x, (1,) * 6).unfold(2, 3, 1).unfold(3, 3, 1).unfold(4, 3, 1)
    y_ref = (xu * K).sum(dim=(-3, -2, -1))
    assert torch.allclose(M.sm_block(img, x, W, b), y_ref, rtol=1e-12, atol=1e-12)


def test_identity_stencil():
    """W = 0, b = e_13 (centre tap) -> y = x."""
    shape = (4, 5, 6)
    img = _rand((1, *shape), 13)
    x = _rand((2, 1, *shape), 14)
    W = torch.zeros(27, 1, 3, 3, 3)
    b = torch.zeros(27)
    b[13] = 1.0
    assert torch.equal(M.sm_block(img, x, W, b), x)
    assert torch.equal(M.sm_block_trans(img, x, W, b), x)


def test_constant_stencil_is_convolution():
    """W = 0 -> K_p = b_p everywhere -> y = cross-correlation of x with b as 3x3x3 kernel;
    the transpose is the convolution with the flipped kernel."""
    shape = (5, 5, 6)
    img = _rand((1, *shape), 15)
    x = _rand((2, 1, *shape), 16)
    W = torch.zeros(27, 1, 3, 3, 3)
    b = _rand((27,), 17)
    y = M.sm_block(img, x, W, b)
    y_ref = F.conv3d(x, b.reshape(1, 1, 3, 3, 3), padding=1)
    assert torch.allclose(y, y_ref, rtol=1e-12, atol=1e-12)
    yt = M.sm_block_trans(img, x, W, b)
    yt_ref = F.conv3d(x, b.flip(0).reshape(1, 1, 3, 3, 3), padding=1)
    assert torch.allclose(yt, yt_ref, rtol=1e-12, atol=1e-12)


def test_tap_order():
    """Only tap p = 9*(d1+1)+3*(d2+1)+(d3+1) with d = (1, 0, -1) nonzero: y[v] = x[v+d]."""
    shape = (4, 4, 4)
    img = torch.zeros(1, *shape)
    x = _rand((1, 1, *shape), 18)
    W = torch.zeros(27, 1, 3, 3, 3)
    b = torch.zeros(27)
    p = 9 * 2 + 3 * 1 + 0
    b[p] = 1.0
    y = M.sm_block(img, x, W, b)
    assert torch.equal(y[0, 0, :-1, :, 1:], x[0, 0, 1:, :, :-1])
    assert torch.all(y[0, 0, -1] == 0) and torch.all(y[0, 0, :, :, 0] == 0)


def test_weight_index_order():
    """W[p, s, e] multiplies img[s, v+e]: a single nonzero weight shifts the image."""
    shape = (4, 5, 6)
    img = _rand((2, *shape), 19)
    x = torch.ones(1, 1, *shape)
    W = torch.zeros(27, 2, 3, 3, 3)
    b = torch.zeros(27)
    W[13, 1, 0, 2, 1] = 1.0  # centre tap, channel 1, e = (-1, +1, 0)
    K = M.stencil_weights(img, W, b)
    assert torch.equal(K[13, 1:, :-1, :], img[1, :-1, 1:, :])
    assert torch.all(K[13, 0] == 0)


@pytest.mark.parametrize("C", [1, 3])
def test_adjoint_identity(C):
    shape = (6, 5, 4)
    img = _rand((C, *shape), 20)
    x = _rand((2, 1, *shape), 21)
    z = _rand((2, 1, *shape), 22)
    W = _rand((27, C, 3, 3, 3), 23)
    b = _rand((27,), 24)
    lhs = (z * M.sm_block(img, x, W, b)).sum()
    rhs = (M.sm_block_trans(img, z, W, b) * x).sum()
    assert abs(float(lhs - rhs)) < 1e-12 * max(1.0, abs(float(lhs)))


@pytest.mark.parametrize("C", [1, 3])
def test_sm_linear(C):
    shape = (3, 4, 5)
    img = _rand((C, *shape), 25)
    w = _rand((1, C, 3, 3, 3), 26)
    b = _rand((1,), 27)
    z, ybar = M.sm_linear(img, w, b)
    zb, ybarb = M.brute_sm_linear(img, w, b)
    assert torch.allclose(ybar, ybarb, rtol=1e-12, atol=1e-14)
    assert abs(float(z) - zb) < 1e-12
    # closed form: mean of the full conv output
    z_conv = F.conv3d(img[None], w, b, padding=1).mean()
    assert abs(float(z - z_conv)) < 1e-12


def test_sm_linear_pyform_mapping():
    """Conv3d(C->27).mean() (sm_model_3d_py.py:27-31) equals the channel-mean weight form
    of sm_model_3d.py:149-153."""
    C, shape = 3, (4, 4, 6)
    img = _rand((C, *shape), 28)
    W27 = _rand((27, C, 3, 3, 3), 29)
    b27 = _rand((27,), 30)
    z_py = M.sm_linear_pyform(img, W27, b27)
    z, _ = M.sm_linear(img, W27.mean(0, keepdim=True), b27.mean(0, keepdim=True))
    assert abs(float(z_py - z)) < 1e-12


def test_avg_pool_and_upsample_1d_cases():
    x = torch.arange(8.0).reshape(1, 1, 2, 2, 2)
    assert float(M.avg_pool(x)) == 3.5
    # trilinear x2 on a 1-D ramp [0, 1] along the last axis: [0, .25, .75, 1]
    r = torch.tensor([0.0, 1.0]).reshape(1, 1, 1, 1, 2).expand(1, 1, 2, 2, 2).contiguous()
    up = M.upsample(r, "trilinear")
    assert torch.allclose(up[0, 0, 0, 0], torch.tensor([0.0, 0.25, 0.75, 1.0]))
    up_n = M.upsample(r, "nearest")
    assert torch.allclose(up_n[0, 0, 0, 0], torch.tensor([0.0, 0.0, 1.0, 1.0]))


@pytest.mark.parametrize("nc", [1, 2, 3, 5])
def test_upsample_matrix_matches_interpolate(nc):
    U = M.upsample_1d_weights(nc)
    x = _rand((nc,), 31)
    y = F.interpolate(x.reshape(1, 1, nc), scale_factor=2, mode="linear")
    assert torch.allclose(U @ x, y.reshape(-1), atol=1e-14)
    assert torch.allclose(U.sum(1), torch.ones(2 * nc))


def test_upsample_separable_trilinear():
    shape = (3, 2, 4)
    x = _rand((1, 1, *shape), 32)
    U1, U2, U3 = (M.upsample_1d_weights(n) for n in shape)
    y = torch.einsum("ai,bj,ck,ijk->abc", U1, U2, U3, x[0, 0])
    assert torch.allclose(M.upsample(x)[0, 0], y, atol=1e-13)


def _brute_model(P, n, image, b, mode, mask, swap):
    """Composition of sm_model_3d.py:264-287 written with the brute-force block."""
    def blk(img, x, name, m=False):
        y = M.brute_sm_block(img, x, P[f"{name}.weight"], P[f"{name}.bias"])
        if m:
            y = torch.where(img[1] == 0, torch.zeros(()), y)
        return y

    def lin(img, name):
        return M.brute_sm_linear(img, P[f"{name}.weight"], P[f"{name}.bias"])[0]

    x = [blk(image, b, "pre.0", mask)]
    imgs = [image]
    for i in range(1, n):
        x.append(M.avg_pool(x[-1]))
        imgs.append(M.avg_pool(imgs[-1]))
        x[-1] = blk(imgs[-1], x[-1], f"pre.{i}")
    x.append(M.avg_pool(x[-1]))
    imgs.append(M.avg_pool(imgs[-1]))
    x[-1] = blk(imgs[-1], x[-1], "l")
    for i in range(n, 0, -1):
        x[i] = M.upsample(x[i], mode)
        c0, c1 = lin(imgs[i - 1], f"c0.{i-1}"), lin(imgs[i - 1], f"c1.{i-1}")
        if swap:
            x[i - 1] = blk(imgs[i - 1], c0 * x[i - 1] + c1 * x[i], f"post.{i-1}", mask and i == 1)
        else:
            x[i] = blk(imgs[i - 1], x[i], f"post.{i-1}", mask and i == 1)
            x[i - 1] = c0 * x[i - 1] + c1 * x[i]
    return x[0]


@pytest.mark.parametrize("mode,mask,swap", [("trilinear", False, False), ("nearest", True, True)])
def test_model_composition_vs_brute(mode, mask, swap):
    n, C, shape = 2, 3, (4, 4, 4)
    g = torch.Generator().manual_seed(33)
    P = M.init_params(n, C, g)
    image = (_rand((C, *shape), 34) > 0).double()
    b = _rand((1, 1, *shape), 35)
    y = M.model_forward(P, n, image, b, mode, mask, swap)
    yb = _brute_model(P, n, image, b, mode, mask, swap)
    assert torch.allclose(y, yb, rtol=1e-11, atol=1e-12)


def test_model_is_linear_in_b():
    n, C, shape = 3, 1, (8, 8, 8)
    P = M.init_params(n, C, torch.Generator().manual_seed(36))
    image = _rand((C, *shape), 37).abs()
    b1, b2 = _rand((1, 1, *shape), 38), _rand((1, 1, *shape), 39)
    y = M.model_forward(P, n, image, 2.0 * b1 - 3.0 * b2)
    y_lin = 2.0 * M.model_forward(P, n, image, b1) - 3.0 * M.model_forward(P, n, image, b2)
    assert torch.allclose(y, y_lin, rtol=1e-11, atol=1e-12)


def test_spd_models_are_symmetric():
    """SPDSMModelDn3D and SmallSPDSMModelDn3D are symmetric linear maps (the property
    sm_model_3d.py:474-486 prints)."""
    n, C, shape = 2, 3, (8, 8, 8)
    image = (_rand((C, *shape), 40) > -0.3).double()
    x, y = _rand((1, 1, *shape), 41), _rand((1, 1, *shape), 42)
    for reduced in (False, True):
        P = M.init_params_spd(n, C, torch.Generator().manual_seed(43), reduced=reduced)
        f = M.small_spd_forward if reduced else M.spd_forward
        a = float((x * f(P, n, image, y)).sum())
        c = float((y * f(P, n, image, x)).sum())
        assert abs(a - c) < 1e-11 * max(1.0, abs(a))


def test_init_bounds():
    """kaiming_uniform(a=sqrt(5)) on (27, C, 3, 3, 3) gives U(+-1/sqrt(27C)) (model.py:22-29)."""
    import torch.nn as nn
    C = 3
    w = torch.empty(27, C, 3, 3, 3)
    nn.init.kaiming_uniform_(w, a=math.sqrt(5))
    assert float(w.abs().max()) <= 1 / math.sqrt(27 * C) + 1e-12
