"""Training of the SM preconditioner on synthetic scenes.

Loss of the benchmark (configured form): mean over the batch of ||b - A P(b)||^2 / ||b||^2
with A the fluid Laplacian applied matrix-free; the reference's residual_loss
(loss_functions.py:18-23, mean of ||b - A P(b)||) is available as well.  The step is
forward -> loss -> backward (hand-written HIP backward kernels behind the autograd
Functions of sm_model_3d, stencils recomputed in backward) -> Adam, as train.py:61-72.

    python -m mlpcg_amd.train --out mlpcg_amd/weights/sm_l5_c1.pt --steps 3000
"""
from __future__ import annotations

import argparse
import json
import math
import os
import time

import numpy as np
import torch

from . import kernels as K
from . import scenes as S
from .sm_model_3d import SmallSMModelDn3D, SmallSPDSMModelDn3D, SPDSMModelDn3D

MODELS = {"sm": lambda n, c: SmallSMModelDn3D(n, c, "trilinear"),
          "smallspd": lambda n, c: SmallSPDSMModelDn3D(n, c),
          "spd": lambda n, c: SPDSMModelDn3D(n, c)}


class LaplacianFunction(torch.autograd.Function):
    """y = M A M x on the full grid, M the restriction to the fluid cells (the matrix-free
    kernel reads neighbours as if x vanished off the fluid, so the input is masked);
    the map is symmetric, so the backward is the same operator.  x: (B, V) float32."""

    @staticmethod
    def forward(ctx, x, code):
        ctx.save_for_backward(code)
        m = (code != 0).reshape(1, -1)
        return torch.stack([K.spmv_flags(code, xi.contiguous()) for xi in x * m])

    @staticmethod
    def backward(ctx, g):
        code, = ctx.saved_tensors
        m = (code != 0).reshape(1, -1)
        return torch.stack([K.spmv_flags(code, gi.contiguous()) for gi in g * m]), None


class SceneBatch:
    """Device-resident scene data: flags code, fluid mask, image channels."""

    def __init__(self, scene: S.Scene, num_imgs: int, device, image_dtype=torch.float32):
        self.scene = scene
        self.shape = scene.shape
        self.flags = torch.as_tensor(scene.flags.astype(np.int8), device=device)
        self.code = K.flags_code(self.flags)
        self.fluid = (self.flags == S.FLUID)
        self.image = torch.as_tensor(S.image_channels(scene.flags, num_imgs), device=device).to(image_dtype)
        lab, closed = S.fluid_components_without_air(scene.flags)
        self.closed = [torch.as_tensor((lab == c).ravel(), device=device) for c in closed]
        self._caches = {}

    def model_input(self, model, dtype):
        """What model.forward_cached needs for this scene (pyramid, mean patches), built on
        first use per (levels, activation type)."""
        key = (model.n, dtype)
        if key not in self._caches:
            self._caches[key] = model.image_cache(self.image, dtype)
        return self._caches[key]


def white_rhs(sb: SceneBatch, count: int, gen: torch.Generator):
    """N(0,1) on fluid cells, mean-subtracted on air-free components, unit norm."""
    V = sb.fluid.numel()
    b = torch.randn(count, V, generator=gen, device=sb.fluid.device) * sb.fluid.reshape(1, V)
    return _finish_rhs(sb, b)


def multiscale_rhs(sb: SceneBatch, count: int, gen: torch.Generator, levels: int = 5):
    """Sum of white noise drawn on the grid and on 2^l-coarser grids, upsampled
    (trilinear); gives every wavelength a comparable share of the energy, which is what
    the reference's low-mode-weighted Ritz right-hand sides provide (preprocess.py:38-58)."""
    n1, n2, n3 = sb.shape
    dev = sb.fluid.device
    b = torch.zeros(count, 1, n1, n2, n3, device=dev)
    for l in range(levels + 1):
        f = 2 ** l
        if n1 % f or n2 % f or n3 % f:
            break
        z = torch.randn(count, 1, n1 // f, n2 // f, n3 // f, generator=gen, device=dev)
        for _ in range(l):
            z = K.upsample2(z.contiguous(), "trilinear")
        b += z
    V = n1 * n2 * n3
    b = b.reshape(count, V) * sb.fluid.reshape(1, V)
    return _finish_rhs(sb, b)


def _finish_rhs(sb, b):
    for m in sb.closed:
        mm = m.reshape(1, -1)
        mean = (b * mm).sum(1, keepdim=True) / mm.sum()
        b = b - mean * mm
    return b / b.norm(dim=1, keepdim=True).clamp(min=1e-30)


def relative_residual_sq(sb: SceneBatch, x, b):
    """mean_b ||b - A x||^2 / ||b||^2 with x restricted to the fluid cells."""
    B = b.shape[0]
    V = b.shape[1]
    xf = x.reshape(B, V).float() * sb.fluid.reshape(1, V)
    r = b - LaplacianFunction.apply(xf, sb.code)
    return ((r * r).sum(1) / (b * b).sum(1)).mean()


class RelResidualSqFunction(torch.autograd.Function):
    """mean_b ||b_b - M A M x_b||^2 / ||b_b||^2 (M: restriction to the fluid cells) with the
    reductions done by the library's two-kernel FP64 dots instead of torch reductions, so a
    training step built on it can be captured in a HIP graph.  A is symmetric, hence
    d/dx_b = -2 / (B ||b_b||^2) M A M r_b.  For FP32 and BF16 activations each term is two
    library kernels (residual with its sums; gradient in the activation type), the same
    arithmetic as x.float() * fluid, the matrix-free A, b - A x, (c * A(r * fluid)) * fluid."""

    @staticmethod
    def forward(ctx, x, b, code, fluid):
        B, V = b.shape
        b = b.contiguous()
        ctx.xshape, ctx.xdtype = x.shape, x.dtype
        ctx.fused = x.dtype in (torch.float32, torch.bfloat16)
        if ctx.fused:
            xs = x.reshape(B, V)
            terms = [K.loss_resid(code, xs[i], b[i]) for i in range(B)]
            rs = [t[0] for t in terms]
            nr = torch.stack([t[1][0] for t in terms])
            nb = torch.stack([t[1][1] for t in terms])
        else:
            xf = (x.reshape(B, V).float() * fluid).contiguous()
            rs = [bi - K.spmv_flags(code, xi) for bi, xi in zip(b, xf)]
            nr = torch.stack([K.dot(ri, ri) for ri in rs])
            nb = torch.stack([K.dot(bi, bi) for bi in b])
        loss = (nr / nb).sum()  # B entries: one workgroup, no partials to clear
        ctx.save_for_backward(nb, code, fluid, *rs)
        return (loss / B).float()

    @staticmethod
    def backward(ctx, g):
        nb, code, fluid, *rs = ctx.saved_tensors
        B = len(rs)
        coef = (-2.0 / B) * g.double() / nb  # (B,)
        if ctx.fused:
            gs = [K.loss_grad(code, ri, c, ctx.xdtype) for c, ri in zip(coef, rs)]
        else:
            gs = [((c.float() * K.spmv_flags(code, (ri * fluid).contiguous())) * fluid).to(ctx.xdtype)
                  for c, ri in zip(coef, rs)]
        gx = gs[0] if B == 1 else torch.stack(gs)
        return gx.reshape(ctx.xshape), None, None, None


def relative_residual_sq_dot(sb: SceneBatch, x, b):
    """relative_residual_sq through RelResidualSqFunction (library reductions)."""
    return RelResidualSqFunction.apply(x, b, sb.code, sb.fluid.reshape(1, -1).float())


def residual_loss(sb: SceneBatch, x, b):
    """loss_functions.py:18-23: mean_b ||b - A x||."""
    B, V = b.shape
    xf = x.reshape(B, V).float() * sb.fluid.reshape(1, V)
    r = b - LaplacianFunction.apply(xf, sb.code)
    return r.norm(dim=1).mean()


def _block_params(model):
    """(module, name, parameter) of the SM blocks' weights and biases: the parameters each
    scene of a step uses once (the SMLinear weights reach the scenes through one shared
    stack)."""
    mods = [m for attr in ("pre", "post") for m in getattr(model, attr, [])]
    if hasattr(model, "l"):
        mods.append(model.l)
    return [(m, n, m._parameters[n]) for m in mods for n in ("weight", "bias") if m._parameters.get(n) is not None]


def _step_body(model, batches, rhs, act_dtype, loss_fn):
    """Forward over the scenes, then one backward of the mean loss; returns the mean loss
    as a device tensor.  The scenes share one stack of the SMLinear weights.  Scenes after
    the first run the SM blocks on aliases of their parameters (same storage, separate
    gradients), so no parameter receives one gradient per scene added one kernel at a time:
    the aliases' gradients join the parameters' in one foreach sum per scene.  With BF16
    activations the blocks' weights are packed for the matrix cores once per step."""
    cached = hasattr(model, "forward_cached")
    blocks = _block_params(model) if cached else []
    if cached and act_dtype == torch.bfloat16:
        # the matrix-core blocks' weight packing, once per step instead of per call
        mods = list(dict.fromkeys(m for m, _, _ in blocks))
        with K.packed_weights([(m.weight, m.bias) for m in mods]):
            return _scenes_step(model, batches, rhs, act_dtype, loss_fn, cached, blocks)
    return _scenes_step(model, batches, rhs, act_dtype, loss_fn, cached, blocks)


def _scenes_step(model, batches, rhs, act_dtype, loss_fn, cached, blocks):
    total = None
    lin = None
    aliases = []
    for k, (sb, b) in enumerate(zip(batches, rhs)):
        bb = b.reshape(b.shape[0], 1, *sb.shape).to(act_dtype)
        if cached:  # the scene image is fixed: reuse its pyramid
            cache = sb.model_input(model, act_dtype)
            if lin is None and hasattr(model, "linear_stack"):
                lin = model.linear_stack(cache[2].dtype)
            al = [p.detach().requires_grad_() for _, _, p in blocks] if k > 0 else []
            for (m, n, _), a in zip(blocks, al):
                m._parameters[n] = a
            try:
                x = model.forward_cached(cache, bb, lin) if lin is not None else model.forward_cached(cache, bb)
            finally:
                for m, n, p in blocks:
                    m._parameters[n] = p
            aliases.append(al)
        else:
            x = model(sb.image.to(act_dtype), bb)
        loss = loss_fn(sb, x, b) / len(batches)
        total = loss if total is None else total + loss
    total.backward()
    for al in aliases[1:]:
        dst, src = [], []
        for (_, _, p), a in zip(blocks, al):
            if a.grad is None:
                continue
            if p.grad is None:
                p.grad = a.grad
            else:
                dst.append(p.grad)
                src.append(a.grad)
        if dst:
            torch._foreach_add_(dst, src)
    return total.detach()


def train_step(model, batches, rhs, optimizer, act_dtype=torch.float32, loss_fn=relative_residual_sq):
    """One Adam step over a list of (SceneBatch, rhs[B, V]) pairs; returns the mean loss."""
    optimizer.zero_grad(set_to_none=True)
    total = _step_body(model, batches, rhs, act_dtype, loss_fn)
    optimizer.step()
    return float(total)


class GraphedTrainStep:
    """train_step captured once as a HIP graph and replayed: each call copies the
    right-hand sides into static buffers and replays forward, backward and the Adam update;
    the loss stays on the device.  The optimizer must be built with capturable=True (Adam
    then keeps its step count on the device).  The loss uses the library's reductions
    (relative_residual_sq_dot), so no torch reduction over many workgroups is recorded.
    The warm-up steps that precede the capture run on copies of `rhs` and are undone
    afterwards (parameters restored, optimizer moments and step counts cleared), so the
    first replay is the optimizer's first step."""

    def __init__(self, model, batches, rhs, optimizer, act_dtype=torch.float32,
                 loss_fn=relative_residual_sq_dot, warmup=2):
        dev = batches[0].code.device
        self.rhs = [b.detach().clone() for b in rhs]
        for sb in batches:  # image pyramids are built outside the graph
            if hasattr(model, "forward_cached"):
                sb.model_input(model, act_dtype)
        params = [p for p in model.parameters()]
        saved = [p.detach().clone() for p in params]
        side = torch.cuda.Stream(device=dev)
        side.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(side):
            for _ in range(warmup):
                optimizer.zero_grad(set_to_none=False)
                _step_body(model, batches, self.rhs, act_dtype, loss_fn)
                optimizer.step()
        torch.cuda.current_stream(dev).wait_stream(side)
        # gradients unset at the capture: each parameter's first gradient of a replay is
        # written in place of the old one (no fill, no accumulate into zeros)
        optimizer.zero_grad(set_to_none=True)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.loss = _step_body(model, batches, self.rhs, act_dtype, loss_fn)
            optimizer.step()
        with torch.no_grad():  # capture ran nothing: undo the warm-up steps
            for p, s0 in zip(params, saved):
                p.copy_(s0)
            for state in optimizer.state.values():
                for v in state.values():
                    if torch.is_tensor(v):
                        v.zero_()

    def __call__(self, rhs):
        """One Adam step on the given right-hand sides; returns the (device) mean loss."""
        for dst, src in zip(self.rhs, rhs):
            dst.copy_(src)
        self.graph.replay()
        return self.loss


def make_scenes(n, seeds, kind="spheres"):
    out = []
    for s in seeds:
        if kind == "spheres":
            out.append(S.scene_spheres(n, s))
        elif kind == "pillars":
            out.append(S.scene_pillars(n, s))
        elif kind == "spheres40":
            out.append(S.scene_for_config(4, s, n))
        else:
            raise ValueError(kind)
    return out


def evaluate_pcg(model, sb: SceneBatch, dtype=torch.float32, method="npcg", max_it=300, tol=1e-6, seed=0):
    from .solver import FusedPCG
    plan = model.make_preconditioner(sb.image.float(), dtype)
    solver = FusedPCG(sb.scene.flags, "flags", device=sb.flags.device)
    b = S.make_rhs(sb.scene.flags, seed)[0]
    x, info = solver.solve(solver.vector(b), plan, method=method, max_it=max_it, tol=tol)
    return info


def evaluate_methods(model, sb: SceneBatch, methods=("npcg", "npcg_flex", "npsd"), **kw):
    out = {}
    for m in methods:
        info = evaluate_pcg(model, sb, method=m, **kw)
        out[m] = (info["iters"], info["rel_res"])
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--levels", type=int, default=5)
    ap.add_argument("--num-imgs", type=int, default=1)
    ap.add_argument("--scenes", type=int, default=16)
    ap.add_argument("--rhs-per-scene", type=int, default=2)
    ap.add_argument("--scenes-per-step", type=int, default=2)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--eval-every", type=int, default=250)
    ap.add_argument("--max-minutes", type=float, default=15.0)
    ap.add_argument("--init", default=None)
    ap.add_argument("--loss", choices=("relsq", "resid"), default="relsq")
    ap.add_argument("--rhs", choices=("mix", "white", "multiscale", "ritz"), default="mix",
                    help="ritz: combinations of each scene's Ritz vectors (preprocess.py:38-58)")
    ap.add_argument("--ritz-vectors", type=int, default=200,
                    help="Lanczos steps per scene for --rhs ritz (the reference uses 1600)")
    ap.add_argument("--select", default="npcg", help="method whose iteration count selects the kept weights")
    ap.add_argument("--model", choices=tuple(MODELS), default="sm")
    ap.add_argument("--act-dtype", choices=("fp32", "bf16"), default="fp32",
                    help="activation/stencil storage in forward and backward (weights stay FP32)")
    ap.add_argument("--kind", choices=("spheres", "pillars", "spheres40"), default="spheres")
    args = ap.parse_args(argv)
    loss_fn = relative_residual_sq if args.loss == "relsq" else residual_loss

    dev = torch.device("cuda:0")
    torch.manual_seed(args.seed)
    model = MODELS[args.model](args.levels, args.num_imgs).to(dev)
    if args.init:
        model.load_state_dict(torch.load(args.init, map_location=dev, weights_only=True))
    opt = torch.optim.Adam(model.parameters(), lr=args.lr)
    sched = torch.optim.lr_scheduler.LambdaLR(
        opt, lambda s: 0.5 * (1 + math.cos(math.pi * min(s, args.steps) / args.steps)) * 0.99 + 0.01)
    act = torch.bfloat16 if args.act_dtype == "bf16" else torch.float32
    scenes = [SceneBatch(s, args.num_imgs, dev)
              for s in make_scenes(args.n, range(100, 100 + args.scenes), args.kind)]
    held = SceneBatch(make_scenes(args.n, [7], args.kind)[0], args.num_imgs, dev)
    gen = torch.Generator(device=dev).manual_seed(args.seed)
    ritz = {}
    if args.rhs == "ritz":  # per-scene Ritz bases, as lib/create_dataset.py prepares them offline
        from . import ritz as RZ
        for i, sb in enumerate(scenes):
            A = S.build_matrix(sb.scene.flags)
            b0 = S.make_rhs(sb.scene.flags, 100 + i)[0]
            _, Rv = RZ.ritz_vectors(RZ.CSRMatrix64(A, dev), b0, args.ritz_vectors)
            ritz[id(sb)] = (Rv, torch.as_tensor(sb.scene.fluid_cells, device=dev))
            print(f"scene {i}: {Rv.shape[0]} Ritz vectors", flush=True)
    t0 = time.time()
    log = []
    best = None
    for step in range(1, args.steps + 1):
        idx = [(step * args.scenes_per_step + k) % len(scenes) for k in range(args.scenes_per_step)]
        bs = [scenes[i] for i in idx]
        rhs = []
        for sb in bs:
            if args.rhs == "ritz":
                from . import ritz as RZ
                Rv, fidx = ritz[id(sb)]
                b = RZ.training_rhs(Rv, args.rhs_per_scene, num_requested=args.ritz_vectors, generator=gen)
                rhs.append(RZ.full_grid(b, fidx, sb.fluid.numel()))
                continue
            h = {"mix": args.rhs_per_scene // 2, "white": 0, "multiscale": args.rhs_per_scene}[args.rhs]
            parts = []
            if args.rhs_per_scene - h:
                parts.append(white_rhs(sb, args.rhs_per_scene - h, gen))
            if h:
                parts.append(multiscale_rhs(sb, h, gen))
            rhs.append(torch.cat(parts))
        loss = train_step(model, bs, rhs, opt, act_dtype=act, loss_fn=loss_fn)
        sched.step()
        if step % 25 == 0:
            print(f"step {step} loss {loss:.5f} lr {sched.get_last_lr()[0]:.2e} t {time.time() - t0:.1f}s",
                  flush=True)
        if step % args.eval_every == 0 or step == args.steps:
            with torch.no_grad():
                ev = evaluate_methods(model, held)
            rec = {"step": step, "loss": loss, "eval": ev, "time_s": time.time() - t0}
            log.append(rec)
            print(json.dumps(rec), flush=True)
            it, res = ev[args.select]
            if res <= 1e-6 and (best is None or it < best[0]):
                best = (it, {k: v.detach().cpu().clone() for k, v in model.state_dict().items()})
            os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
            sd = best[1] if best is not None else {k: v.detach().cpu() for k, v in model.state_dict().items()}
            torch.save(sd, args.out)
        if (time.time() - t0) / 60 > args.max_minutes:
            print("time limit reached", flush=True)
            break
    with open(os.path.splitext(args.out)[0] + ".json", "w") as f:
        json.dump({"args": vars(args), "log": log, "best_iters": best[0] if best else None}, f, indent=1)


if __name__ == "__main__":
    main()
