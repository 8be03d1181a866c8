"""GPU training path: the configured loss, its gradient through the HIP backward kernels,
BF16 activations with FP32 master weights, and a few Adam steps."""
import numpy as np
import pytest
import torch

from mlpcg_amd import scenes as S

pytestmark = pytest.mark.gpu


def _setup(dev, n=32, seed=0, levels=3):
    from mlpcg_amd import train as T
    from mlpcg_amd.sm_model_3d import SmallSMModelDn3D
    torch.manual_seed(seed)
    model = SmallSMModelDn3D(levels, 1).to(dev)
    sb = T.SceneBatch(S.scene_spheres(n, seed, num_spheres=3, rmin=3, rmax=6), 1, dev)
    gen = torch.Generator(device=dev).manual_seed(seed)
    return T, model, sb, gen


def test_laplacian_function_is_symmetric(dev):
    T, _, sb, gen = _setup(dev)
    x = T.white_rhs(sb, 2, gen).requires_grad_()
    y = T.LaplacianFunction.apply(x, sb.code)
    g = torch.randn_like(y)
    (y * g).sum().backward()
    ref = T.LaplacianFunction.apply(g * sb.fluid.reshape(1, -1), sb.code)
    assert torch.allclose(x.grad, ref)


def test_loss_matches_host_formula(dev):
    T, model, sb, gen = _setup(dev)
    b = T.white_rhs(sb, 2, gen)
    with torch.no_grad():
        x = model(sb.image, b.reshape(2, 1, *sb.shape))
        loss = float(T.relative_residual_sq(sb, x, b))
    A = S.build_matrix(sb.scene.flags)
    f = sb.scene.fluid_cells
    xs = x.reshape(2, -1).double().cpu().numpy()[:, f]
    bs = b.double().cpu().numpy()[:, f]
    r = bs - (A @ xs.T).T
    ref = float(np.mean((r * r).sum(1) / (bs * bs).sum(1)))
    assert abs(loss - ref) < 1e-5 * max(1.0, ref)


def test_whole_model_gradient_finite_difference_f64(dev):
    """d/dtheta <R, P(b)> along a random parameter direction for the whole FP64 model
    (every block, SMLinear, pool and upsample backward), central differences."""
    T, model, sb, gen = _setup(dev, n=16, levels=2)
    model = model.double()
    img = sb.image.double()
    b = T.white_rhs(sb, 2, gen).double()
    R = torch.randn(2, 1, *sb.shape, dtype=torch.float64, device=dev)
    model.zero_grad()
    (model(img, b.reshape(2, 1, *sb.shape)) * R).sum().backward()
    params = [p for p in model.parameters()]
    dirs = [torch.randn_like(p) for p in params]
    an = sum(float((p.grad * d).sum()) for p, d in zip(params, dirs))
    h = 1e-5
    with torch.no_grad():
        for p, d in zip(params, dirs):
            p.add_(h * d)
        lp = float((model(img, b.reshape(2, 1, *sb.shape)) * R).sum())
        for p, d in zip(params, dirs):
            p.sub_(2 * h * d)
        lm = float((model(img, b.reshape(2, 1, *sb.shape)) * R).sum())
        for p, d in zip(params, dirs):
            p.add_(h * d)
    fd = (lp - lm) / (2 * h)
    assert abs(fd - an) <= 1e-6 * max(1.0, abs(an))


@pytest.mark.parametrize("act", [torch.float32, torch.bfloat16])
def test_adam_steps_reduce_loss(dev, act):
    T, model, sb, gen = _setup(dev)
    opt = torch.optim.Adam(model.parameters(), lr=3e-3)
    b = T.white_rhs(sb, 4, gen)
    with torch.no_grad():
        l0 = float(T.relative_residual_sq(sb, model(sb.image, b.reshape(4, 1, *sb.shape)), b))
    for _ in range(30):
        T.train_step(model, [sb], [b], opt, act_dtype=act)
    with torch.no_grad():
        l1 = float(T.relative_residual_sq(sb, model(sb.image, b.reshape(4, 1, *sb.shape)), b))
    assert l1 < 0.7 * l0
    for p in model.parameters():
        assert p.dtype == torch.float32 and torch.isfinite(p).all()


def test_bf16_gradients_close_to_fp32(dev):
    T, model, sb, gen = _setup(dev)
    b = T.white_rhs(sb, 2, gen)
    grads = {}
    for act in (torch.float32, torch.bfloat16):
        model.zero_grad()
        x = model(sb.image.to(act), b.reshape(2, 1, *sb.shape).to(act))
        T.relative_residual_sq(sb, x, b).backward()
        grads[act] = torch.cat([p.grad.flatten() for p in model.parameters()])
    g32, g16 = grads[torch.float32], grads[torch.bfloat16]
    assert float((g16 - g32).norm() / g32.norm()) < 5e-2
