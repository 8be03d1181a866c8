"""FP64 CPU restatement of the SM preconditioner (test infrastructure only).

Index conventions (shared with the HIP kernels):
  grid (N1, N2, N3), last axis fastest; tap p = 9*(d1+1) + 3*(d2+1) + (d3+1) for
  d in {-1,0,1}^3 (sm_block_3d_kernel.cu:46-52); W is (27, C, 3, 3, 3) with
  W[p, s, e] multiplying img[s, v + e] (sm_block_3d_kernel.cu:54-61); values outside
  the grid are zero.
"""
from __future__ import annotations

import itertools
import math

import torch
import torch.nn.functional as F

TAPS = [(d1, d2, d3) for d1 in (-1, 0, 1) for d2 in (-1, 0, 1) for d3 in (-1, 0, 1)]


def _shifted(x, d):
    """out[..., v] = x[..., v + d] with zeros outside; x has 3 trailing grid dims."""
    xp = F.pad(x, (1, 1, 1, 1, 1, 1))
    n1, n2, n3 = x.shape[-3:]
    return xp[..., 1 + d[0]:1 + d[0] + n1, 1 + d[1]:1 + d[1] + n2, 1 + d[2]:1 + d[2] + n3]


def stencil_weights(img, W, b):
    """K_p(v) = b_p + sum_{s,e} W[p,s,e] img[s, v+e]  -> (27, N1, N2, N3).

    sm_model_3d_py.py:16 (Conv3d C->27, padding 'same'); sm_block_3d_kernel.cu:52-64.
    """
    return F.conv3d(img[None], W, b, padding=1)[0]


def sm_block(img, x, W, b, mask=False):
    """y[c, v] = sum_p K_p(v) x[c, v + d_p]   (sm_model_3d_py.py:15-22, A.1).

    img (C, N1, N2, N3); x (B, 1, N1, N2, N3).  mask: y = 0 where img[1] == 0
    (sm_model_3d.py:87-88).
    """
    K = stencil_weights(img, W, b)
    y = torch.zeros_like(x)
    for p, d in enumerate(TAPS):
        y = y + K[p] * _shifted(x, d)
    if mask:
        y = torch.where(img[1] == 0, torch.zeros((), dtype=y.dtype), y)
    return y


def sm_block_trans(img, x, W, b):
    """y[c, v] = sum_p K_p(v - d_p) x[c, v - d_p]  (sm_block_trans_3d_kernel.cu:46-73)."""
    K = stencil_weights(img, W, b)
    y = torch.zeros_like(x)
    for p, d in enumerate(TAPS):
        y = y + _shifted(K[p] * x, tuple(-t for t in d))
    return y


def sm_linear(img, w, b):
    """ybar[s, e] = (1/V) sum_v img[s, v+e];  z = sum w*ybar + b.

    sm_linear_3d_kernel.cu:62-83 (mean patch), :185-186 (z = w . ybar + b).
    Returns (z of shape (1,), ybar of shape (1, C, 3, 3, 3)).
    """
    C = img.shape[0]
    V = img.shape[1] * img.shape[2] * img.shape[3]
    ybar = torch.zeros(1, C, 3, 3, 3, dtype=img.dtype)
    for s in range(C):
        for e in itertools.product((-1, 0, 1), repeat=3):
            ybar[0, s, e[0] + 1, e[1] + 1, e[2] + 1] = _shifted(img[s], e).sum() / V
    z = (w.flatten() * ybar.flatten()).sum().reshape(1) + b
    return z, ybar


def sm_linear_pyform(img, W27, b27):
    """SmallLinearBlock3DPY form: Conv3d(C->27)(img).mean() (sm_model_3d_py.py:24-31).
    Equal to sm_linear(img, W27.mean(0, keepdim=True), b27.mean(0, keepdim=True))."""
    return F.conv3d(img[None], W27, b27, padding=1).mean()


def avg_pool(x):
    """F.avg_pool3d(x, (2,2,2)) on (..., N1, N2, N3)."""
    lead = x.shape[:-3]
    y = F.avg_pool3d(x.reshape(-1, 1, *x.shape[-3:]), (2, 2, 2))
    return y.reshape(*lead, *y.shape[-3:])


def upsample(x, mode="trilinear"):
    """F.interpolate(x, scale_factor=2, mode) with align_corners=False."""
    lead = x.shape[:-3]
    xx = x.reshape(-1, 1, *x.shape[-3:])
    if mode == "nearest":
        y = F.interpolate(xx, scale_factor=2)
    else:
        y = F.interpolate(xx, scale_factor=2, mode=mode)
    return y.reshape(*lead, *y.shape[-3:])


def upsample_1d_weights(nc):
    """Dense (2nc x nc) 1-D linear upsampling matrix with align_corners=False:
    source index max((f + 0.5)/2 - 0.5, 0), upper neighbour clamped to nc-1."""
    U = torch.zeros(2 * nc, nc, dtype=torch.float64)
    for f in range(2 * nc):
        src = max((f + 0.5) / 2 - 0.5, 0.0)
        i0 = int(math.floor(src))
        i1 = min(i0 + 1, nc - 1)
        lam = src - i0
        U[f, i0] += 1 - lam
        U[f, i1] += lam
    return U


# ----------------------------------------------------------------------------
# Model composition (sm_model_3d.py:164-287; Appendix A.5 of the layout notes)
# ----------------------------------------------------------------------------

def init_params(n, C, generator=None, dtype=torch.float64):
    """Seeded parameters with the reference state_dict keys and shapes.

    Block weights use kaiming_uniform(a=sqrt(5)) -> U(+-1/sqrt(fan_in)) with fan_in =
    27*C, bias U(+-1/sqrt(fan_in)) (model.py:22-29).  Linear weights are the mean over
    27 of such a (27,C,3,3,3) draw (sm_model_3d.py:149-153).
    """
    g = generator
    bound = 1.0 / math.sqrt(27 * C)

    def u(*shape):
        return (torch.rand(*shape, generator=g, dtype=torch.float64) * 2 - 1) * bound

    P = {}
    for i in range(n):
        P[f"pre.{i}.weight"] = u(27, C, 3, 3, 3)
        P[f"pre.{i}.bias"] = u(27)
    for i in range(n):
        P[f"post.{i}.weight"] = u(27, C, 3, 3, 3)
        P[f"post.{i}.bias"] = u(27)
    P["l.weight"] = u(27, C, 3, 3, 3)
    P["l.bias"] = u(27)
    for name in ("c0", "c1"):
        for i in range(n):
            P[f"{name}.{i}.weight"] = u(27, C, 3, 3, 3).mean(0, keepdim=True)
            P[f"{name}.{i}.bias"] = u(27).mean(0, keepdim=True)
    return {k: v.to(dtype) for k, v in P.items()}


def model_forward(P, n, image, b, mode="trilinear", mask=False, swap=False):
    """SmallSMModelDn3D.forward (sm_model_3d.py:264-287).  b: (B, 1, N1, N2, N3)."""
    x = [sm_block(image, b, P["pre.0.weight"], P["pre.0.bias"], mask=mask)]
    imgs = [image]
    for i in range(1, n):
        x.append(avg_pool(x[-1]))
        imgs.append(avg_pool(imgs[-1]))
        x[-1] = sm_block(imgs[-1], x[-1], P[f"pre.{i}.weight"], P[f"pre.{i}.bias"])
    x.append(avg_pool(x[-1]))
    imgs.append(avg_pool(imgs[-1]))
    x[-1] = sm_block(imgs[-1], x[-1], P["l.weight"], P["l.bias"])
    for i in range(n, 0, -1):
        x[i] = upsample(x[i], mode)
        c0, _ = sm_linear(imgs[i - 1], P[f"c0.{i-1}.weight"], P[f"c0.{i-1}.bias"])
        c1, _ = sm_linear(imgs[i - 1], P[f"c1.{i-1}.weight"], P[f"c1.{i-1}.bias"])
        m = mask and (i - 1 == 0)
        Wp, bp = P[f"post.{i-1}.weight"], P[f"post.{i-1}.bias"]
        if swap:
            x[i - 1] = c0 * x[i - 1] + c1 * x[i]
            x[i - 1] = sm_block(imgs[i - 1], x[i - 1], Wp, bp, mask=m)
        else:
            x[i] = sm_block(imgs[i - 1], x[i], Wp, bp, mask=m)
            x[i - 1] = c0 * x[i - 1] + c1 * x[i]
    return x[0]


def init_params_spd(n, C=3, generator=None, dtype=torch.float64, reduced=False):
    """Parameters of SPDSMModelDn3D (l0, l1, l, c0, c1; sm_model_3d.py:293-311) or, with
    reduced=True, SmallSPDSMModelDn3D (l0, l, c0, c1; :388-402)."""
    g = generator
    bound = 1.0 / math.sqrt(27 * C)

    def u(*shape):
        return (torch.rand(*shape, generator=g, dtype=torch.float64) * 2 - 1) * bound

    P = {}
    for i in range(n):
        P[f"l0.{i}.weight"] = u(27, C, 3, 3, 3)
        P[f"l0.{i}.bias"] = u(27)
        if not reduced:
            P[f"l1.{i}.weight"] = u(27, C, 3, 3, 3)
            P[f"l1.{i}.bias"] = u(27)
    P["l.weight"] = u(27, C, 3, 3, 3)
    P["l.bias"] = u(27)
    for name in ("c0", "c1"):
        for i in range(n):
            P[f"{name}.{i}.weight"] = u(27, C, 3, 3, 3).mean(0, keepdim=True)
            P[f"{name}.{i}.bias"] = u(27).mean(0, keepdim=True)
    return {k: v.to(dtype) for k, v in P.items()}


def spd_forward(P, n, image, b):
    """SPDSMModelDn3D.forward (sm_model_3d.py:349-383); nearest upsampling / 8."""
    x = [sm_block(image, b, P["l0.0.weight"], P["l0.0.bias"])]
    imgs = [image]
    c0, c1 = [], []
    for i in range(1, n):
        x.append(avg_pool(x[-1]))
        imgs.append(avg_pool(imgs[-1]))
        x[-1] = sm_block(imgs[-1], x[-1], P[f"l0.{i}.weight"], P[f"l0.{i}.bias"])
    x.append(avg_pool(x[-1]))
    imgs.append(avg_pool(imgs[-1]))
    x[-1] = sm_block(imgs[-1], x[-1], P["l.weight"], P["l.bias"])
    for i in range(n, 0, -1):
        x[i] = upsample(x[i], "nearest") / 8
        c0.insert(0, sm_linear(imgs[i - 1], P[f"c0.{i-1}.weight"], P[f"c0.{i-1}.bias"])[0])
        c1.insert(0, sm_linear(imgs[i - 1], P[f"c1.{i-1}.weight"], P[f"c1.{i-1}.bias"])[0])
        x[i] = sm_block(imgs[i - 1], x[i], P[f"l1.{i-1}.weight"], P[f"l1.{i-1}.bias"])
        x[i - 1] = c0[0] * x[i - 1] + c1[0] * x[i]
    bb = x[0]
    x = [bb]
    for i in range(n):
        x.append(sm_block_trans(imgs[i], x[i], P[f"l1.{i}.weight"], P[f"l1.{i}.bias"]))
        x[-1] = avg_pool(x[-1])
    x[-1] = sm_block_trans(imgs[-1], x[-1], P["l.weight"], P["l.bias"])
    for i in range(n - 1, -1, -1):
        x[i + 1] = upsample(x[i + 1], "nearest") / 8
        x[i] = c0[i] * x[i] + c1[i] * x[i + 1]
        x[i] = sm_block_trans(imgs[i], x[i], P[f"l0.{i}.weight"], P[f"l0.{i}.bias"])
    return x[0]


def small_spd_forward(P, n, image, b):
    """SmallSPDSMModelDn3D.forward (sm_model_3d.py:404-423)."""
    x = [sm_block(image, b, P["l0.0.weight"], P["l0.0.bias"])]
    imgs = [image]
    for i in range(1, n):
        x.append(avg_pool(x[-1]))
        imgs.append(avg_pool(imgs[-1]))
        x[-1] = sm_block(imgs[-1], x[-1], P[f"l0.{i}.weight"], P[f"l0.{i}.bias"])
    x.append(avg_pool(x[-1]))
    imgs.append(avg_pool(imgs[-1]))
    x[-1] = sm_block(imgs[-1], x[-1], P["l.weight"], P["l.bias"])
    x[-1] = sm_block_trans(imgs[-1], x[-1], P["l.weight"], P["l.bias"])
    for i in range(n, 0, -1):
        x[i] = upsample(x[i], "nearest") / 8
        c0 = sm_linear(imgs[i - 1], P[f"c0.{i-1}.weight"], P[f"c0.{i-1}.bias"])[0]
        c1 = sm_linear(imgs[i - 1], P[f"c1.{i-1}.weight"], P[f"c1.{i-1}.bias"])[0]
        x[i - 1] = c0 ** 2 * x[i - 1] + c1 ** 2 * x[i]
        x[i - 1] = sm_block_trans(imgs[i - 1], x[i - 1], P[f"l0.{i-1}.weight"], P[f"l0.{i-1}.bias"])
    return x[0]


# ----------------------------------------------------------------------------
# Pure-Python brute force (tiny grids only): the defining sums, written out.
# ----------------------------------------------------------------------------

def _at(a, idx):
    for t, n in zip(idx, a.shape[-3:]):
        if t < 0 or t >= n:
            return 0.0
    return float(a[..., idx[0], idx[1], idx[2]])


def brute_K(img, W, b, v):
    C = img.shape[0]
    out = []
    for p in range(27):
        z = float(b[p])
        for s in range(C):
            for e in itertools.product((-1, 0, 1), repeat=3):
                u = (v[0] + e[0], v[1] + e[1], v[2] + e[2])
                z += float(W[p, s, e[0] + 1, e[1] + 1, e[2] + 1]) * _at(img[s], u)
        out.append(z)
    return out


def brute_sm_block(img, x, W, b):
    B = x.shape[0]
    N = img.shape[1:]
    y = torch.zeros_like(x)
    for v in itertools.product(*(range(n) for n in N)):
        K = brute_K(img, W, b, v)
        for c in range(B):
            acc = 0.0
            for p, d in enumerate(TAPS):
                acc += K[p] * _at(x[c, 0], (v[0] + d[0], v[1] + d[1], v[2] + d[2]))
            y[c, 0, v[0], v[1], v[2]] = acc
    return y


def brute_sm_block_trans(img, x, W, b):
    B = x.shape[0]
    N = img.shape[1:]
    y = torch.zeros_like(x)
    for v in itertools.product(*(range(n) for n in N)):
        for c in range(B):
            acc = 0.0
            for p, d in enumerate(TAPS):
                u = (v[0] - d[0], v[1] - d[1], v[2] - d[2])
                xv = _at(x[c, 0], u)
                if xv != 0.0:
                    acc += brute_K(img, W, b, u)[p] * xv
            y[c, 0, v[0], v[1], v[2]] = acc
    return y


def brute_sm_linear(img, w, b):
    C = img.shape[0]
    N = img.shape[1:]
    V = N[0] * N[1] * N[2]
    ybar = torch.zeros(1, C, 3, 3, 3, dtype=torch.float64)
    for s in range(C):
        for e in itertools.product((-1, 0, 1), repeat=3):
            acc = 0.0
            for v in itertools.product(*(range(n) for n in N)):
                acc += _at(img[s], (v[0] + e[0], v[1] + e[1], v[2] + e[2]))
            ybar[0, s, e[0] + 1, e[1] + 1, e[2] + 1] = acc / V
    z = float((w.flatten().double() * ybar.flatten()).sum()) + float(b[0])
    return z, ybar
