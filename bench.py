#!/usr/bin/env python
"""Benchmark: PCG time-to-1e-6 relative residual at 256^3 with the BF16 vs the FP32
SM preconditioner, plus Adam train steps/s of the preconditioner.

One step = one full solve of the 256^3 pillars scene (24 vertical solid pillars of
radius 4-10, solid walls, air above z = 0.55 N) with the l=5 SmallSMModelDn3D
preconditioner (C=1 image diag(A)/6): the image-dependent setup (pyramid, stencils,
SMLinear coefficients) followed by npsd (or --method) to ||b - Ax||/||b|| <= 1e-6, the
residual recomputed in FP64 every iteration.  `value` is the mean seconds per step with
the BF16 preconditioner; inputs are resident in HBM before timing.

roofline is that of the dominant kernel of the BF16 solve, the fused matrix-core
level-0 block; roofline_streamed that of the stored-stencil level-0 apply (used by the
FP32 plan).  Their traffic fields are the HBM bytes per launch read by two rocprofv3
counter passes (FETCH_SIZE, WRITE_SIZE, calibrated on a 1 GiB copy), recorded in
profiles/sm_fused_bf16_level0_pillars256_hbm_traffic.txt and
profiles/sm_apply_bf16_level0_pillars256_hbm_traffic.txt for the 256^3 workload only
(null otherwise).

three_channel: the same scene solved with the three-channel model (image channels
[air, fluid, non-solid], mlpcg_amd/weights/sm_l5_c3.pt) by the BF16 fused plan, the same
with the non-solid channel folded into the other two in the fused blocks, the BF16
stored-stencil plan and the FP32 plan (--skip-c3 leaves it out).

train: Adam steps/s of the graph-captured training step on 8 x 128^3 scenes (FP32), and
train.train256 that of the 4 x 256^3 training step (40-sphere scenes, shipped weights,
right-hand sides from a pool resident in HBM) with BF16 and with FP32 activations, and
their ratio, for the one-channel model and (three_channel) the three-channel one
(--skip-train leaves both out).

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "PCG time-to-1e-6 rel. residual at 256³, BF16 vs FP32 precond; train steps/s"
HBM_PEAK_GBS = 8000.0
DEFAULT_WEIGHTS = os.path.join(ROOT, "mlpcg_amd", "weights", "sm_l5_c1.pt")
C3_WEIGHTS = os.path.join(ROOT, "mlpcg_amd", "weights", "sm_l5_c3.pt")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


TRAFFIC_FUSED = os.path.join(ROOT, "profiles", "sm_fused_bf16_level0_pillars256_hbm_traffic.txt")
TRAFFIC_STREAMED = os.path.join(ROOT, "profiles", "sm_apply_bf16_level0_pillars256_hbm_traffic.txt")


def recorded_traffic(n, path):
    """HBM bytes per launch of a level-0 kernel at 256^3 from the recorded counter passes."""
    if n != 256 or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch"):
            return int(line.split("=")[1])
    return None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--levels", type=int, default=5)
    ap.add_argument("--method", default="npsd",
                    help="solver of the benchmark value (npsd: the reference's default, test.py:327)")
    ap.add_argument("--also", default="npcg_flex", help="comma list of further methods reported")
    ap.add_argument("--max-it", type=int, default=300)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--weights", default=DEFAULT_WEIGHTS)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-train", action="store_true")
    ap.add_argument("--skip-fp16", action="store_true")
    ap.add_argument("--skip-c3", action="store_true", help="skip the three-channel model solves")
    ap.add_argument("--roofline-iters", type=int, default=20)
    ap.add_argument("--profile-only", action="store_true",
                    help="run only the dominant kernel (for rocprofv3 counter passes)")
    ap.add_argument("--workload", default="pillars256",
                    choices=("pillars256", "spheres128", "train256", "sphere64cpu"),
                    help="pillars256: the benchmark line; the others print their own report")
    ap.add_argument("--train-steps", type=int, default=300)
    ap.add_argument("--skip-heldout", action="store_true",
                    help="train256: skip the held-out solves (kernel traces of the training steps alone)")
    ap.add_argument("--graphs", type=int, default=0, help="replay solver iterations as HIP graphs")
    return ap.parse_args()


def sync():
    import torch
    torch.cuda.synchronize()


def timed(fn):
    sync()
    t0 = time.perf_counter()
    out = fn()
    sync()
    return time.perf_counter() - t0, out


def load_model(path, levels, dev, num_imgs=1):
    import torch
    from mlpcg_amd.sm_model_3d import SmallSMModelDn3D
    model = SmallSMModelDn3D(levels, num_imgs, "trilinear").to(dev)
    source = "random-init"
    if path and os.path.exists(path):
        model.load_state_dict(torch.load(path, map_location=dev, weights_only=True))
        source = os.path.relpath(path, ROOT)
    else:
        torch.manual_seed(0)
        model = SmallSMModelDn3D(levels, num_imgs, "trilinear").to(dev)
    return model.eval(), source


def three_channel_solves(args, sc, solver, b, dev):
    """The same scene and RHS with a three-channel model ([air, fluid, non-solid] images,
    the reference's own models): BF16 with the fused matrix-core apply, the same with the
    non-solid channel folded into the other two, BF16 with stored stencils, FP32; npsd to
    the same tolerance."""
    import torch
    from mlpcg_amd import scenes as S
    model, wsrc = load_model(C3_WEIGHTS, args.levels, dev, num_imgs=3)
    image = torch.as_tensor(S.image_channels(sc.flags, 3), device=dev)
    out = {"weights": wsrc}
    # bf16_fused_folded: the fused blocks take [air, fluid] with the non-solid channel's
    # weights added to theirs (non-solid = air + fluid on these images)
    for name, dtype, fused, fold in (("bf16_fused", torch.bfloat16, 1, False),
                                     ("bf16_fused_folded", torch.bfloat16, 1, True),
                                     ("bf16_stored", torch.bfloat16, 0, False), ("fp32", torch.float32, 0, False)):
        plan = model.make_preconditioner(image, dtype, fused_mode=fused, fold_third=fold)

        def step():
            plan.update(image)
            return solver.solve(b, plan, method=args.method, max_it=args.max_it, tol=args.tol)
        for _ in range(args.warmup):
            step()
        times, info = [], None
        for _ in range(args.steps):
            t, (_, info) = timed(step)
            times.append(t)
        out[name] = {"s_per_solve": round(float(np.mean(times)), 6), "iters": info["iters"],
                     "status": info["status"], "rel_res_device_fp64": info["rel_res"],
                     "plan_mib": round(plan.nbytes / 2 ** 20, 1), "fused_blocks": plan.fused_blocks}
        log(f"[bench] three channels {name}: {out[name]['s_per_solve'] * 1e3:.2f} ms/solve, "
            f"{info['iters']} iterations, status {info['status']}, plan {out[name]['plan_mib']} MiB")
        del plan
    return out


def roofline_apply(model, image, r, dtype, iters):
    """Level-0 SM block apply (first block of the preconditioner: FP32 input scaled by
    1/||r||, K and output in `dtype`), timed with HIP events on torch's current stream,
    which is the stream the library launches on."""
    import torch
    from mlpcg_amd import kernels as K
    W, b = model.pre[0].weight, model.pre[0].bias
    Kd = K.stencil_weights(image, W, b, dtype=dtype)
    scale = torch.tensor(1.0 / float(r.norm()), dtype=torch.float64, device=r.device)
    x = r.reshape(1, 1, *r.shape[-3:]).contiguous()
    for _ in range(3):
        K.sm_apply(Kd, x, out_dtype=dtype, in_scale=scale)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(iters):
        K.sm_apply(Kd, x, out_dtype=dtype, in_scale=scale)
    e1.record()
    e1.synchronize()
    t = e0.elapsed_time(e1) / 1e3 / iters
    V = x.numel()
    esz = torch.tensor([], dtype=dtype).element_size()
    alg = 27 * esz * V + 4 * V + esz * V
    return t, alg


# K contraction 2*27*(27C+1) (bias as the constant input) + apply 2*27, C = 1
FUSED_FLOP_PER_VOXEL = 2 * 27 * 28 + 2 * 27


def roofline_fused(plan, r, iters):
    """Level-0 pre-smoother block of a BF16 plan alone (block 0: FP32 input scaled by
    1/||r||, BF16 output; the fused matrix-core kernel), timed with HIP events on torch's
    current stream, which is the stream the library launches on.  Algorithmic bytes per
    voxel: x 4 + FP16 image 2 + y 2."""
    import torch
    scale = torch.tensor(1.0 / float(r.norm()), dtype=torch.float64, device=r.device)
    x = r.reshape(r.shape[-3:]).contiguous()
    for _ in range(3):
        plan.apply_block(0, x, torch.bfloat16, scale)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(iters):
        plan.apply_block(0, x, torch.bfloat16, scale)
    e1.record()
    e1.synchronize()
    t = e0.elapsed_time(e1) / 1e3 / iters
    V = x.numel()
    return t, 8 * V, FUSED_FLOP_PER_VOXEL * V


def copy_bandwidth(dev, nbytes=2 << 30):
    import torch
    a = torch.empty(nbytes // 4, dtype=torch.float32, device=dev).fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(2):
        b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        b.copy_(a)
    e1.record()
    e1.synchronize()
    t = e0.elapsed_time(e1) / 1e3 / 10
    return 2 * nbytes / t / 1e9


def train_rate(dev, steps=3, warmup=1, n=128, scenes=8, levels=5, graphed_steps=20):
    """Adam steps/s on 8 seeded 128^3 scenes (one N(0,1) RHS each), loss
    ||b - A P(b)||^2/||b||^2, hand-written backward kernels: eager steps, and the same step
    captured in a HIP graph (right-hand sides from a pool of 4 per scene drawn before the
    timed loop, copied in before every replay)."""
    import torch
    from mlpcg_amd import train as T
    from mlpcg_amd.sm_model_3d import SmallSMModelDn3D
    batches = [T.SceneBatch(s, 1, dev) for s in T.make_scenes(n, range(200, 200 + scenes))]
    gen = torch.Generator(device=dev).manual_seed(2)
    rhs = [T.white_rhs(sb, 1, gen) for sb in batches]
    torch.manual_seed(1)
    model = SmallSMModelDn3D(levels, 1, "trilinear").to(dev)
    opt = torch.optim.Adam(model.parameters(), lr=1e-4)
    for _ in range(warmup):
        T.train_step(model, batches, rhs, opt)
    t, _ = timed(lambda: [T.train_step(model, batches, rhs, opt) for _ in range(steps)])
    eager = steps / t
    torch.manual_seed(1)
    model = SmallSMModelDn3D(levels, 1, "trilinear").to(dev)
    opt = torch.optim.Adam(model.parameters(), lr=1e-4, capturable=True, fused=True)
    step = T.GraphedTrainStep(model, batches, rhs, opt)
    pool = [[T.white_rhs(sb, 1, gen) for sb in batches] for _ in range(4)]
    step(pool[0])
    t, _ = timed(lambda: [step(pool[k % 4]) for k in range(graphed_steps)])
    loss = float(step.loss)
    return {"steps_per_s": round(graphed_steps / t, 3), "eager_steps_per_s": round(eager, 3),
            "last_loss": loss, "peak_hbm_GiB": round(torch.cuda.max_memory_allocated(dev) / 2 ** 30, 2)}


def precond_probe(flags):
    """The preconditioner input of the precision check: N(0,1) on the fluid cells (seed 0),
    unit norm, on the full grid (float32, host)."""
    from mlpcg_amd import scenes as S
    fluid = flags.ravel() == S.FLUID
    r = np.random.default_rng(0).standard_normal(int(fluid.sum()))
    g = np.zeros(flags.size, dtype=np.float32)
    g[fluid] = r / np.linalg.norm(r)
    return g.reshape(flags.shape)


def cpu_baseline(flags, image_np, model, iters, levels, plan_outputs=None):
    """The oracle (FP32 torch on the host cores) for one 256^3 preconditioner apply plus
    the per-iteration SpMVs and dots with scipy, extrapolated to `iters` iterations.  As
    the checker, the same leg also runs the oracle in FP64 on the same input and returns
    the relative error of each GPU plan's output (plan_outputs: name -> host tensor)."""
    import torch
    import scipy.sparse  # noqa: F401
    from mlpcg_amd import scenes as S
    from oracle import model as M
    cores = len(os.sched_getaffinity(0))
    torch.set_num_threads(cores)
    P = {k: v.detach().float().cpu() for k, v in model.state_dict().items()}
    img = torch.as_tensor(image_np, dtype=torch.float32)
    A = S.build_matrix(flags)
    g = precond_probe(flags)
    r = g.ravel()[flags.ravel() == S.FLUID].astype(np.float64)
    x = torch.as_tensor(g).reshape(1, 1, *flags.shape)
    with torch.no_grad():
        t_apply, _ = _wall(lambda: M.model_forward(P, levels, img, x))
    t_spmv, _ = _wall(lambda: (A @ r, A @ r))
    t_dots, _ = _wall(lambda: (r @ r, r @ r, r @ r))
    per_it = t_apply + t_spmv + t_dots
    out = {"value": per_it * iters, "unit": "s", "cores": cores, "kind": "port",
           "sample": (f"oracle FP32 torch preconditioner apply at 256^3 ({t_apply:.2f} s) + 2 scipy "
                      f"SpMV + 3 dots ({t_spmv + t_dots:.3f} s) on {cores} host threads, "
                      f"x {iters} iterations (extrapolated, FP32-preconditioner iteration count)"),
           "per_iteration_s": per_it}
    if plan_outputs:
        P64 = {k: v.double() for k, v in P.items()}
        with torch.no_grad():
            z64 = M.model_forward(P64, levels, img.double(), x.double()).reshape(flags.shape)
        nz = float(z64.norm())
        out["oracle_fp64_rel_err"] = {k: float((z.double().reshape(flags.shape) - z64).norm() / nz)
                                      for k, z in plan_outputs.items()}
    return out


def _wall(fn):
    t0 = time.perf_counter()
    out = fn()
    return time.perf_counter() - t0, out


def workload_spheres128(args, dev):
    """128^3, walls + 6 seeded spheres (r 6-20), seeded water level; l=5 FP32
    preconditioner; the CSR SpMV and the matrix-free SpMV solves side by side."""
    import torch
    from mlpcg_amd import scenes as S
    from mlpcg_amd.solver import FusedPCG
    sc = S.scene_spheres(128, args.seed)
    model, wsrc = load_model(args.weights, args.levels, dev)
    image = torch.as_tensor(S.image_channels(sc.flags, 1), device=dev)
    A = S.build_matrix(sc.flags)
    b_np = S.make_rhs(sc.flags, args.seed)[0]
    out = {"workload": "128^3 six spheres, l=5 FP32 preconditioner", "weights": wsrc, "method": args.method}
    plan = model.make_preconditioner(image, torch.float32)
    for layout in ("flags", "csr"):
        s = FusedPCG(sc.flags, layout, A=A if layout == "csr" else None, device=dev,
                     graphs=bool(args.graphs))
        b = s.vector(b_np)
        for precond, tag in ((plan, "mlpcg"), (None, "cg")):
            s.solve(b, precond, method=args.method, max_it=2000, tol=args.tol)
            ts, infos = [], []
            for _ in range(args.steps):
                t, (x, info) = timed(lambda: s.solve(b, precond, method=args.method, max_it=2000, tol=args.tol))
                ts.append(t)
                infos.append(info)
            xs = x.cpu().numpy().astype(np.float64)
            if layout == "flags":
                xs = xs[sc.fluid_cells]
            res = float(np.linalg.norm(b_np - A @ xs) / np.linalg.norm(b_np))
            out[f"{layout}_{tag}"] = {"s_per_solve": float(np.mean(ts)), "iters": infos[-1]["iters"],
                                      "rel_res_host_fp64": res}
            log(f"[bench] spheres128 {layout} {tag}: {np.mean(ts) * 1e3:.2f} ms, {infos[-1]['iters']} it, res {res:.2e}")
    print(json.dumps(out), flush=True)


def train256_scenes(dev, num_imgs=1):
    from mlpcg_amd import train as T
    return [T.SceneBatch(s, num_imgs, dev) for s in T.make_scenes(256, range(300, 304), "spheres40")]


TRAIN256_POOL = 16  # right-hand sides per scene drawn before the timed loop


def train256_steps(args, dev, scenes, act_name, steps, gen, weights=None, num_imgs=1):
    """`steps` Adam steps on the 4 training scenes at 256^3 from the shipped weights, each
    step one HIP-graph replay with BF16 or FP32 activations (FP32 master weights).  The
    right-hand sides are a dataset resident in HBM, as the reference's training set is
    precomputed (lib/create_dataset.py) rather than drawn per step: TRAIN256_POOL white-noise
    right-hand sides per scene are drawn before the timed loop and cycled, and each step
    copies its four into the graph's input buffers (timed).  -> (model, steps/s, last loss,
    peak HBM GiB of the step without the pool, weights source)."""
    import torch
    from mlpcg_amd import train as T
    act = {"bf16": torch.bfloat16, "fp32": torch.float32}[act_name]
    model, wsrc = load_model(weights or args.weights, args.levels, dev, num_imgs=num_imgs)
    model.train()
    opt = torch.optim.Adam(model.parameters(), lr=3e-4, capturable=True, fused=True)
    pool = [[T.white_rhs(sb, 1, gen) for sb in scenes] for _ in range(min(steps, TRAIN256_POOL))]
    pool_bytes = sum(b.numel() * b.element_size() for rhs in pool for b in rhs)
    torch.cuda.reset_peak_memory_stats(dev)
    step = T.GraphedTrainStep(model, scenes, pool[0], opt, act_dtype=act)  # one HIP graph per step
    sync()
    t0 = time.perf_counter()
    for k in range(steps):
        loss_t = step(pool[k % len(pool)])
        if k % 25 == 0:
            log(f"[bench] train256 {act_name} step {k} loss {float(loss_t):.4f}")
    sync()
    sps = steps / (time.perf_counter() - t0)
    peak = (torch.cuda.max_memory_allocated(dev) - pool_bytes) / 2 ** 30
    return model, sps, float(step.loss), peak, wsrc


def train256_rate(args, dev, steps=40):
    """The 256^3 training step of the train256 workload, BF16 vs FP32 activations
    (steps/s after the graph capture, no held-out solves), with the one-channel shipped
    weights and (three_channel, unless --skip-c3) the three-channel ones."""
    import torch
    scenes = train256_scenes(dev)
    gen = torch.Generator(device=dev).manual_seed(3)
    out = {"workload": "4 scenes x 256^3 (40 spheres), 1 RHS each, l=5, Adam from the shipped weights; "
                       "BF16 vs FP32 activations (FP32 master weights); step captured in a HIP graph; "
                       f"right-hand sides from a resident pool of {TRAIN256_POOL} per scene",
           "steps": steps}
    for name in ("bf16", "fp32"):
        _, sps, loss, peak, _ = train256_steps(args, dev, scenes, name, steps, gen)
        out[name] = {"steps_per_s": round(sps, 3), "peak_hbm_GiB": round(peak, 2), "last_loss": loss}
    out["bf16_vs_fp32"] = round(out["bf16"]["steps_per_s"] / out["fp32"]["steps_per_s"], 3)
    if not args.skip_c3 and os.path.exists(C3_WEIGHTS):
        # the reference's three-channel model ([air, fluid, non-solid] images) on the same scenes
        del scenes
        scenes3 = train256_scenes(dev, 3)
        c3 = {"weights": os.path.relpath(C3_WEIGHTS, ROOT)}
        for name in ("bf16", "fp32"):
            _, sps, loss, peak, _ = train256_steps(args, dev, scenes3, name, steps, gen, C3_WEIGHTS, 3)
            c3[name] = {"steps_per_s": round(sps, 3), "peak_hbm_GiB": round(peak, 2), "last_loss": loss}
        c3["bf16_vs_fp32"] = round(c3["bf16"]["steps_per_s"] / c3["fp32"]["steps_per_s"], 3)
        out["three_channel"] = c3
    return out


def workload_train256(args, dev):
    """Training at 256^3 (4 scenes, 40 spheres r 4-30, air fraction 20-50 %): BF16
    forward/backward with FP32 master weights vs FP32, stencils recomputed in backward,
    each Adam step replayed as one HIP graph (fresh right-hand sides every step);
    steps/s, peak HBM and PCG iterations on 8 held-out 256^3 scenes after the same
    number of steps from the same starting weights."""
    import torch
    from mlpcg_amd import train as T
    from mlpcg_amd.solver import FusedPCG
    from mlpcg_amd import scenes as S
    scenes = train256_scenes(dev)
    held = [T.SceneBatch(s, 1, dev) for s in T.make_scenes(256, range(400, 408), "spheres40")]
    out = {"workload": "train 4 x 256^3 (40 spheres), l=5", "steps": args.train_steps}
    gen = torch.Generator(device=dev).manual_seed(3)
    for act_name in ("bf16", "fp32"):
        model, sps, loss, peak, wsrc = train256_steps(args, dev, scenes, act_name, args.train_steps, gen)
        its = []
        model.eval()
        with torch.no_grad():
            for sb in ([] if args.skip_heldout else held):
                plan = model.make_preconditioner(sb.image, torch.float32)
                s = FusedPCG(sb.scene.flags, "flags", device=dev)
                _, info = s.solve(s.vector(S.make_rhs(sb.scene.flags, 0)[0]), plan, method=args.method,
                                  max_it=args.max_it, tol=args.tol)
                its.append(info["iters"] if info["status"] == 0 else None)
        out[act_name] = {"steps_per_s": round(sps, 3), "peak_hbm_GiB": round(peak, 2), "last_loss": loss,
                         "heldout_iters": its, "start_weights": wsrc}
        log(f"[bench] train256 {act_name}: {sps:.2f} steps/s, peak {peak:.1f} GiB, held-out iters {its}")
    print(json.dumps(out), flush=True)


def workload_sphere64cpu(args):
    """64^3 sphere scene, l=4, seeded random FP64 weights, C=1: the oracle's npcg on the
    host CPU in FP64 (<= 200 iterations or 1e-6)."""
    import torch
    from mlpcg_amd import scenes as S
    from oracle import model as M
    from oracle import solvers as O
    cores = len(os.sched_getaffinity(0))
    torch.set_num_threads(cores)
    sc = S.scene_sphere64()
    A = S.build_matrix(sc.flags)
    b = S.make_rhs(sc.flags, 0)[0]
    P = M.init_params(4, 1, torch.Generator().manual_seed(0))
    img = torch.as_tensor(S.image_channels(sc.flags, 1)).double()
    predict = O.make_predict(lambda x: M.model_forward(P, 4, img, x), sc.fluid_cells, sc.shape)
    t, (x, k, hist) = _wall(lambda: O.npcg(b, A, np.zeros_like(b), predict, 200, tol=1e-6))
    out = {"workload": "64^3 sphere, l=4 random-init FP64, CPU oracle npcg", "seconds": t, "iters": k,
           "final_rel_res": float(np.linalg.norm(b - A @ x) / np.linalg.norm(b)), "cores": cores}
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: the preconditioned solve of one scene is a "
              "single-GPU workload here (its z-slab decomposition over GPUs is a design, not "
              "a code path)", file=sys.stderr)
        sys.exit(2)
    import torch
    from mlpcg_amd import scenes as S
    from mlpcg_amd.solver import FusedPCG

    if args.workload == "sphere64cpu":
        workload_sphere64cpu(args)
        return
    dev = torch.device("cuda:0")
    torch.cuda.init()
    if args.workload == "spheres128":
        workload_spheres128(args, dev)
        return
    if args.workload == "train256":
        workload_train256(args, dev)
        return
    log(f"[bench] scene {args.n}^3 pillars seed {args.seed}")
    sc = S.scene_pillars(args.n, args.seed)
    model, wsrc = load_model(args.weights, args.levels, dev)
    image_np = S.image_channels(sc.flags, 1)
    image = torch.as_tensor(image_np, device=dev)
    b_np = S.make_rhs(sc.flags, args.seed)[0]
    solver = FusedPCG(sc.flags, "flags", device=dev, graphs=bool(args.graphs))
    b = solver.vector(b_np)
    r0 = b.reshape(sc.shape)

    if args.profile_only:
        # the dominant kernel (fused level-0 block), the stored-stencil apply, then a plain
        # 1 GiB device copy whose counter readings calibrate FETCH_SIZE / WRITE_SIZE
        plan = model.make_preconditioner(image, torch.bfloat16)
        t, alg, _ = roofline_fused(plan, r0, args.roofline_iters)
        log(f"[bench] profile-only: sm_fused bf16 {t * 1e3:.3f} ms, {alg / t / 1e9:.1f} GB/s, "
            f"algorithmic bytes {alg}")
        del plan
        t, alg = roofline_apply(model, image, r0, torch.bfloat16, args.roofline_iters)
        log(f"[bench] profile-only: sm_apply bf16 {t * 1e3:.3f} ms, {alg / t / 1e9:.1f} GB/s, "
            f"algorithmic bytes {alg}")
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        sync()
        log("[bench] profile-only: copy kernel 1 GiB read + 1 GiB write done")
        return

    plans = {"bf16": model.make_preconditioner(image, torch.bfloat16),
             "fp32": model.make_preconditioner(image, torch.float32)}
    if not args.skip_fp16:
        plans["fp16"] = model.make_preconditioner(image, torch.float16)

    def step(name, method):
        plans[name].update(image)
        return solver.solve(b, plans[name], method=method, max_it=args.max_it, tol=args.tol)

    def run_method(method, names):
        out = {}
        for name in names:
            for _ in range(args.warmup):
                step(name, method)
            k = args.steps if name in ("bf16", "fp32") else max(1, args.steps // 2)
            times, infos = [], []
            xs = None
            for _ in range(k):
                t, (xs, info) = timed(lambda: step(name, method))
                times.append(t)
                infos.append(info)
            t_setup, _ = timed(lambda: plans[name].update(image))
            out[name] = {"s_per_solve": float(np.mean(times)), "min_s": float(np.min(times)),
                         "iters": infos[-1]["iters"], "status": infos[-1]["status"],
                         "rel_res_device_fp64": infos[-1]["rel_res"], "setup_s": t_setup, "x": xs}
            log(f"[bench] {method} {name}: {out[name]['s_per_solve'] * 1e3:.2f} ms/solve, "
                f"{out[name]['iters']} iterations, status {out[name]['status']}")
        return out

    results = run_method(args.method, list(plans))
    c3 = None
    if not args.skip_c3 and os.path.exists(C3_WEIGHTS):
        c3 = three_channel_solves(args, sc, solver, b, dev)
    others = {}
    for m in [t for t in args.also.split(",") if t and t != args.method]:
        o = run_method(m, [n for n in ("bf16", "fp32") if n in plans])
        for v in o.values():
            v.pop("x")
        others[m] = {k: {kk: vv for kk, vv in v.items()} for k, v in o.items()}

    # FP64 check of the returned iterates on the host
    A = S.build_matrix(sc.flags)
    fluid = sc.fluid_cells
    nb = np.linalg.norm(b_np)
    for name, rres in results.items():
        x = rres.pop("x").cpu().numpy()[fluid].astype(np.float64)
        rres["rel_res_host_fp64"] = float(np.linalg.norm(b_np - A @ x) / nb)

    # every plan on the precision probe; the FP64 oracle output it is compared with comes
    # from the CPU leg (the FP64 library plan stands in when that leg is skipped)
    probe = torch.as_tensor(precond_probe(sc.flags), device=dev)
    plan_out = {name: p.apply(probe).cpu() for name, p in plans.items()}
    if args.skip_cpu:
        plan64 = model.make_preconditioner(image, torch.float64)
        z64 = plan64.apply(probe.double()).cpu()
        for name, z in plan_out.items():
            results[name]["precond_rel_err_vs_fp64_plan"] = float((z.double() - z64).norm() / z64.norm())
        del plan64, z64

    # roofline of the dominant kernel: the fused level-0 block of the BF16 plan.  Its
    # arithmetic intensity (FUSED_FLOP_PER_VOXEL / 8 B = 196 FLOP/B) is below the
    # BF16/FP16 ridge point (2.5 PFLOP/s / 8 TB/s = 312), so its roofline is HBM.
    t_k, alg, flop = roofline_fused(plans["bf16"], r0, args.roofline_iters)
    achieved = alg / t_k / 1e9
    copy_gbs = copy_bandwidth(dev)
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": recorded_traffic(args.n, TRAFFIC_FUSED),
            "kernel": "sm_fused_kernel<f32 x, bf16 y, C=1> level 0 (pre.0)",
            "avg_ms": round(t_k * 1e3, 4), "algorithmic_bytes": alg,
            "algorithmic_tflops": round(flop / t_k / 1e12, 1),
            "mfma_frac": round(flop / t_k / 2.5e15, 4),
            "measured_copy_GBs": round(copy_gbs, 1)}
    log(f"[bench] roofline fused: {t_k * 1e6:.1f} us, {achieved:.0f} GB/s ({achieved / HBM_PEAK_GBS:.1%} of HBM), "
        f"{flop / t_k / 1e12:.0f} TFLOP/s")
    t_s, alg_s = roofline_apply(model, image, r0, torch.bfloat16, args.roofline_iters)
    ach_s = alg_s / t_s / 1e9
    roof_streamed = {"bound": "hbm", "achieved": round(ach_s, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach_s / HBM_PEAK_GBS, 4),
                     "traffic": recorded_traffic(args.n, TRAFFIC_STREAMED),
                     "kernel": "sm_apply_kernel<bf16,f32,bf16,bf16,VZ=8> level 0 (stored BF16 stencils)",
                     "avg_ms": round(t_s * 1e3, 4), "algorithmic_bytes": alg_s}
    log(f"[bench] roofline streamed: {t_s * 1e6:.1f} us, {ach_s:.0f} GB/s ({ach_s / HBM_PEAK_GBS:.1%} of peak), "
        f"copy {copy_gbs:.0f} GB/s")

    train = None
    if not args.skip_train:
        train = train_rate(dev)
        train["workload"] = "8 scenes x 128^3, 1 RHS each, l=5, FP32, Adam; step captured in a HIP graph"
        log(f"[bench] train: {train['steps_per_s']:.2f} steps/s graphed, "
            f"{train['eager_steps_per_s']:.2f} eager")
        t256 = train256_rate(args, dev)
        train["train256"] = t256
        log(f"[bench] train256: bf16 {t256['bf16']['steps_per_s']:.2f} steps/s, fp32 "
            f"{t256['fp32']['steps_per_s']:.2f} steps/s ({t256['bf16_vs_fp32']:.2f}x)")
        if "three_channel" in t256:
            t3 = t256["three_channel"]
            log(f"[bench] train256 three channels: bf16 {t3['bf16']['steps_per_s']:.2f} steps/s, fp32 "
                f"{t3['fp32']['steps_per_s']:.2f} steps/s ({t3['bf16_vs_fp32']:.2f}x)")

    cpu = None
    if not args.skip_cpu:
        cpu = cpu_baseline(sc.flags, image_np, model, results["fp32"]["iters"], args.levels, plan_out)
        for name, e in cpu.pop("oracle_fp64_rel_err").items():
            results[name]["precond_rel_err_vs_oracle_fp64"] = e
            log(f"[bench] preconditioner {name}: relative error vs the FP64 oracle {e:.2e}")
        log(f"[bench] cpu baseline: {cpu['value']:.1f} s (extrapolated)")

    bf, f32 = results["bf16"], results["fp32"]
    out = {
        "metric": METRIC,
        "value": round(bf["s_per_solve"], 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(bf["s_per_solve"] * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic",
        "config": {"workload": f"{args.n}^3 pillars scene, l={args.levels} SM preconditioner "
                               f"(BF16 storage, FP32 accumulate), {args.method} to {args.tol:g}",
                   "grid": args.n, "levels": args.levels, "method": args.method,
                   "weights": wsrc, "graphs": bool(args.graphs), "parallelism": "single GPU"},
        "roofline": roof,
        "roofline_streamed": roof_streamed,
        "cpu_baseline": cpu,
        "fp32_s_per_solve": round(f32["s_per_solve"], 6),
        "bf16_speedup_vs_fp32": round(f32["s_per_solve"] / bf["s_per_solve"], 3),
        "iters": {k: v["iters"] for k, v in results.items()},
        "iter_ratio_bf16_vs_fp32": round(bf["iters"] / max(f32["iters"], 1), 3),
        "solves": results,
        "other_methods": others,
        "three_channel": c3,
        "train": train,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
