"""Training step captured as a HIP graph: the library-reduction loss against the torch
formula, and graph replays against eager steps (every replay checked, with eager work and
host synchronisations between the replays)."""
import copy

import pytest
import torch

from mlpcg_amd import scenes as S

pytestmark = pytest.mark.gpu


def _setup(dev, n=32, levels=3, scenes=2, B=2):
    from mlpcg_amd import train as T
    from mlpcg_amd.sm_model_3d import SmallSMModelDn3D
    torch.manual_seed(3)
    model = SmallSMModelDn3D(levels, 1).to(dev)
    sbs = [T.SceneBatch(S.scene_spheres(n, 10 + s, num_spheres=3, rmin=3, rmax=6), 1, dev)
           for s in range(scenes)]
    gen = torch.Generator(device=dev).manual_seed(4)
    return T, model, sbs, gen, B


@pytest.mark.parametrize("dt", [torch.float32, torch.bfloat16])
@pytest.mark.parametrize("cdt", [torch.float32, torch.bfloat16])
@pytest.mark.parametrize("n,off", [(1 << 20, 0), (4099, 0), (4104, 1), (0, 0)])
def test_comb_kernels_match_torch(dev, dt, cdt, n, off):
    """The level combination kernels against torch's expression: forward and input
    gradients bit-exact (same roundings), coefficient gradients against FP64 sums of the
    exact products; vector (n % 8 == 0, aligned), ragged and misaligned (offset views)
    lengths and the empty level."""
    from mlpcg_amd import kernels as K
    g = torch.Generator(device=dev).manual_seed(n + off)
    mk = lambda: torch.randn(n + off, device=dev, generator=g).to(dt)[off:]
    x0, x1, gr = mk(), mk(), mk()
    c0 = torch.tensor([0.7318], device=dev).to(cdt)
    c1 = torch.tensor([-1.2371], device=dev).to(cdt)
    y = K.comb(c0, x0, c1, x1)
    assert torch.equal(y, c0.to(dt) * x0 + c1.to(dt) * x1)
    g0, g1, dots = K.comb_backward(gr, c0, x0, c1, x1)
    assert torch.equal(g0, c0.to(dt) * gr) and torch.equal(g1, c1.to(dt) * gr)
    ref = torch.stack([(gr.double() * x0.double()).sum(), (gr.double() * x1.double()).sum()])
    scale = torch.stack([(gr.double() * x0.double()).abs().sum(), (gr.double() * x1.double()).abs().sum()])
    assert bool(((dots - ref).abs() <= 1e-13 * scale + 1e-300).all())
    # either input gradient may be skipped; the sums may be returned rounded to FP32
    h0, h1, d2 = K.comb_backward(gr, c0, x0, c1, x1, need_g0=False)
    assert h0 is None and torch.equal(h1, g1) and torch.equal(d2, dots)
    _, _, d32 = K.comb_backward(gr, c0, x0, c1, x1, out_dtype=torch.float32)
    assert d32.dtype == torch.float32 and torch.equal(d32, dots.float())


@pytest.mark.parametrize("act,gtol", [(torch.float32, 1e-5), (torch.bfloat16, 8e-3)])
def test_dot_loss_matches_torch_loss_and_gradient(dev, act, gtol):
    """The library loss (fused residual and gradient kernels for FP32 / BF16 outputs) against
    the torch composition on the same network output: the loss to FP64-sum order, the
    gradient to FP32 order (FP32) or to the BF16 rounding of the gradient (BF16)."""
    T, model, sbs, gen, B = _setup(dev)
    sb = sbs[0]
    b = T.white_rhs(sb, B, gen)
    x0 = model(sb.image, b.reshape(B, 1, *sb.shape)).detach().to(act)
    vals, grads = [], []
    for fn in (T.relative_residual_sq, T.relative_residual_sq_dot):
        x = x0.clone().requires_grad_()
        loss = fn(sb, x, b)
        loss.backward()
        vals.append(float(loss.detach()))
        grads.append(x.grad.detach().clone().float())
    assert abs(vals[0] - vals[1]) <= 1e-5 * max(1.0, abs(vals[0]))
    assert float((grads[0] - grads[1]).norm() / grads[0].norm()) < gtol


@pytest.mark.parametrize("act,fused", [(torch.float32, False), (torch.bfloat16, False), (torch.bfloat16, True)])
def test_graphed_step_matches_eager_steps(dev, act, fused):
    """fused: torch's fused Adam (one kernel for all parameters), as the benchmarks use."""
    T, model, sbs, gen, B = _setup(dev)
    model_g = copy.deepcopy(model)
    opt_e = torch.optim.Adam(model.parameters(), lr=3e-3, capturable=True, fused=fused)
    opt_g = torch.optim.Adam(model_g.parameters(), lr=3e-3, capturable=True, fused=fused)
    steps = [[T.white_rhs(sb, B, gen) for sb in sbs] for _ in range(4)]
    graphed = T.GraphedTrainStep(model_g, sbs, steps[0], opt_g, act_dtype=act)
    for p, q in zip(model.parameters(), model_g.parameters()):
        assert torch.equal(p, q)  # warm-up undone
    for k, rhs in enumerate(steps):
        le = T.train_step(model, sbs, rhs, opt_e, act_dtype=act, loss_fn=T.relative_residual_sq_dot)
        lg = float(graphed(rhs))
        torch.cuda.synchronize()
        # eager work between replays
        junk = T.LaplacianFunction.apply(rhs[0], sbs[0].code).sum()
        assert torch.isfinite(junk)
        assert abs(le - lg) <= 1e-5 * max(1.0, abs(le)), (k, le, lg)
        for p, q in zip(model.parameters(), model_g.parameters()):
            assert torch.allclose(p, q, rtol=1e-5, atol=1e-6), k
    # replaying again with the first right-hand sides gives the loss an eager step gives
    le = T.train_step(model, sbs, steps[0], opt_e, act_dtype=act, loss_fn=T.relative_residual_sq_dot)
    assert abs(le - float(graphed(steps[0]))) <= 1e-5 * max(1.0, abs(le))


def test_each_replay_matches_an_eager_step_64(dev):
    """64^3 scenes (level-0 reductions span many workgroups): before every replay the
    graphed model's parameters and Adam state are copied into an eager twin, which takes
    one eager step on the same right-hand sides; all parameters must agree after each
    step, including the level-combine coefficients whose gradients are full-grid dots."""
    from mlpcg_amd import train as T
    from mlpcg_amd.sm_model_3d import SmallSMModelDn3D
    torch.manual_seed(5)
    mg = SmallSMModelDn3D(4, 1).to(dev)
    me = copy.deepcopy(mg)
    sbs = [T.SceneBatch(S.scene_spheres(64, 20 + i, num_spheres=3, rmin=4, rmax=10), 1, dev) for i in range(2)]
    gen = torch.Generator(device=dev).manual_seed(6)
    seq = [[T.white_rhs(sb, 1, gen) for sb in sbs] for _ in range(5)]
    og = torch.optim.Adam(mg.parameters(), lr=3e-3, capturable=True)
    oe = torch.optim.Adam(me.parameters(), lr=3e-3, capturable=True)
    step = T.GraphedTrainStep(mg, sbs, seq[0], og)
    names = [n for n, _ in mg.named_parameters()]
    for k, rhs in enumerate(seq):
        with torch.no_grad():
            for p, q in zip(me.parameters(), mg.parameters()):
                p.copy_(q)
        oe.state.clear()
        for pe, pg in zip(me.parameters(), mg.parameters()):
            if pg in og.state:
                oe.state[pe] = {kk: (v.clone() if torch.is_tensor(v) else v) for kk, v in og.state[pg].items()}
        le = T.train_step(me, sbs, rhs, oe, loss_fn=T.relative_residual_sq_dot)
        lg = float(step(rhs))
        torch.cuda.synchronize()
        assert abs(le - lg) <= 1e-6 * max(1.0, abs(le)), (k, le, lg)
        for n, p, q in zip(names, me.parameters(), mg.parameters()):
            assert torch.allclose(p, q, rtol=1e-6, atol=1e-7), (k, n, float((p - q).abs().max()))


def test_graphed_steps_reduce_loss(dev):
    T, model, sbs, gen, B = _setup(dev)
    opt = torch.optim.Adam(model.parameters(), lr=3e-3, capturable=True)
    rhs = [T.white_rhs(sb, 4, gen) for sb in sbs]
    graphed = T.GraphedTrainStep(model, sbs, rhs, opt)
    l0 = float(graphed(rhs))
    for _ in range(30):
        graphed(rhs)
    l1 = float(graphed(rhs))
    assert l1 < 0.7 * l0
