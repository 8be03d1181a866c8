"""Full-size workloads: the 256^3 pillars scene's BF16 / FP16 / FP32 plans against the FP64
oracle with the shipped trained weights, and the 256^3 training step (4 scenes of 40
spheres, BF16 activations with FP32 master weights, stencils recomputed in backward)."""
import os

import numpy as np
import pytest
import torch

from mlpcg_amd import scenes as S
from oracle import model as M

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WEIGHTS = os.path.join(ROOT, "mlpcg_amd", "weights", "sm_l5_c1.pt")


def _trained_model(dev):
    from mlpcg_amd.sm_model_3d import SmallSMModelDn3D
    m = SmallSMModelDn3D(5, 1, "trilinear").to(dev)
    m.load_state_dict(torch.load(WEIGHTS, map_location=dev, weights_only=True))
    return m.eval()


def _probe(flags):
    fluid = flags.ravel() == S.FLUID
    r = np.random.default_rng(0).standard_normal(int(fluid.sum()))
    g = np.zeros(flags.size, dtype=np.float32)
    g[fluid] = r / np.linalg.norm(r)
    return g.reshape(flags.shape)


# Relative (norm-wise) error bounds of P(r) against the FP64 oracle at 256^3.  FP32 is the
# north-star bar; the 16-bit plans store stencils and level vectors with 8 (BF16) or 11
# (FP16) significand bits, accumulate in FP32, and are also bounded well below the
# difference between a trained and an untrained preconditioner.
PLAN_TOL = {torch.float32: 1e-5, torch.float16: 2e-3, torch.bfloat16: 1e-2}


def test_pillars256_plans_vs_fp64_oracle(dev):
    print("pillars 256^3: FP64 oracle on the host, then FP32/FP16/BF16 plans", flush=True)
    torch.set_num_threads(max(1, min(16, len(os.sched_getaffinity(0)))))
    sc = S.scene_pillars(256, 0)
    m = _trained_model(dev)
    img = torch.as_tensor(S.image_channels(sc.flags, 1))
    x = torch.as_tensor(_probe(sc.flags))
    P64 = {k: v.detach().double().cpu() for k, v in m.state_dict().items()}
    with torch.no_grad():
        z64 = M.model_forward(P64, 5, img.double(), x.double().reshape(1, 1, *sc.shape)).reshape(sc.shape)
    errs = {}
    for dt, tol in PLAN_TOL.items():
        plan = m.make_preconditioner(img.to(dev), dt)
        if dt == torch.bfloat16:
            assert plan.fused_blocks >= 2  # levels 0-2 on the matrix-core path
        z = plan.apply(x.to(dev)).cpu().double()
        errs[dt] = float((z - z64).norm() / z64.norm())
        print(f"  {dt}: {errs[dt]:.3e} (bound {tol:g})", flush=True)
        del plan
    for dt, tol in PLAN_TOL.items():
        assert errs[dt] <= tol, (dt, errs[dt])


def _spheres40(dev, seeds):
    from mlpcg_amd import train as T
    return [T.SceneBatch(s, 1, dev) for s in T.make_scenes(256, seeds, "spheres40")]


def _grads(model, sb, b, act):
    from mlpcg_amd import train as T
    model.zero_grad(set_to_none=True)
    x = model.forward_cached(sb.model_input(model, act), b.reshape(b.shape[0], 1, *sb.shape).to(act))
    T.relative_residual_sq(sb, x, b).backward()
    return torch.cat([p.grad.detach().double().flatten() for p in model.parameters()])


def test_train256_bf16_step(dev):
    """The 256^3 training workload: graphed BF16 Adam steps on 4 x 256^3 spheres40 scenes with FP32 master
    weights.  The loss stays finite and falls over 20 steps on fixed right-hand sides; the
    BF16 parameter gradient is within 5e-2 (norm-wise) of the FP32 one at the same
    weights; peak HBM stays under 4 GiB, while keeping the stencils of the level-0 pre and
    post blocks of the four scenes for backward would alone take 7.2 GB."""
    from mlpcg_amd import train as T
    from mlpcg_amd.sm_model_3d import SmallSMModelDn3D
    print("train256: 4 scenes x 256^3, BF16 graphed steps", flush=True)
    torch.manual_seed(5)
    model = SmallSMModelDn3D(5, 1, "trilinear").to(dev)
    scenes = _spheres40(dev, range(300, 304))
    gen = torch.Generator(device=dev).manual_seed(3)
    rhs = [T.white_rhs(sb, 1, gen) for sb in scenes]
    # gradient agreement BF16 vs FP32 at the initial weights (one scene, eager)
    g32 = _grads(model, scenes[0], rhs[0], torch.float32)
    g16 = _grads(model, scenes[0], rhs[0], torch.bfloat16)
    gerr = float((g16 - g32).norm() / g32.norm())
    print(f"  BF16 vs FP32 gradient: {gerr:.3e}", flush=True)
    assert torch.isfinite(g16).all() and gerr <= 5e-2
    model.zero_grad(set_to_none=True)
    for sb in scenes:  # activations of the FP32 gradient check are not part of the step
        sb._caches.pop((model.n, torch.float32), None)
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    torch.cuda.reset_peak_memory_stats(dev)
    opt = torch.optim.Adam(model.parameters(), lr=1e-3, capturable=True)
    step = T.GraphedTrainStep(model, scenes, rhs, opt, act_dtype=torch.bfloat16)
    losses = [float(step(rhs)) for _ in range(20)]
    peak = torch.cuda.max_memory_allocated(dev) / 2 ** 30
    print(f"  losses {losses[0]:.4f} -> {losses[-1]:.4f}, peak {peak:.2f} GiB", flush=True)
    assert all(np.isfinite(losses))
    assert np.mean(losses[-5:]) < 0.9 * np.mean(losses[:5])
    assert peak < 4.0
    for p in model.parameters():
        assert p.dtype == torch.float32 and torch.isfinite(p).all()


def test_bf16_block_backward_vs_oracle_on_spheres40_crop(dev):
    """The BF16 SM block backward (stencils recomputed, transposed gather for dx, matrix-
    core weight gradient) on a 32^3 crop of a spheres40 scene against FP64 autograd of the
    oracle on the same BF16-rounded inputs."""
    from mlpcg_amd import kernels as K
    sc = S.scene_for_config(4, 300, 256)
    img = torch.as_tensor(S.image_channels(sc.flags, 1))[:, 96:128, 96:128, 96:128].contiguous()
    g = torch.Generator().manual_seed(11)
    W = (torch.rand(27, 1, 3, 3, 3, generator=g, dtype=torch.float64) * 2 - 1) / 27 ** 0.5
    b = (torch.rand(27, generator=g, dtype=torch.float64) * 2 - 1) / 27 ** 0.5
    x = torch.randn(2, 1, 32, 32, 32, generator=g, dtype=torch.float64).bfloat16()
    go = torch.randn(2, 1, 32, 32, 32, generator=g, dtype=torch.float64).bfloat16()
    imb = img.bfloat16()
    gx, gW, gb = K.sm_block_backward(go.to(dev), imb.to(dev), x.to(dev), W.float().to(dev), b.float().to(dev))
    xr = x.double().requires_grad_()
    Wr = W.float().double().requires_grad_()
    br = b.float().double().requires_grad_()
    (M.sm_block(imb.double(), xr, Wr, br) * go.double()).sum().backward()
    rel = lambda a, r: float((a.cpu().double() - r).norm() / r.norm())
    e = (rel(gx, xr.grad), rel(gW, Wr.grad), rel(gb, br.grad))
    print(f"  BF16 backward rel errors gx {e[0]:.2e} gW {e[1]:.2e} gb {e[2]:.2e}", flush=True)
    # gx is stored in BF16 (2^-9 relative rounding) from BF16 stencils; gW/gb accumulate
    # in FP32 from BF16 operands
    assert e[0] < 1e-2 and e[1] < 1e-2 and e[2] < 1e-2
