"""GPU parity of the fused matrix-core apply (K recomputed per voxel, fused.hip) for 16-bit
plans: against the stored-K plan on exactly representable data (tile indexing, halos,
taps, channels, masks, epilogue), against the FP64 oracle with trained weights, and in a
solve."""
import os

import numpy as np
import pytest
import torch

from mlpcg_amd import scenes as S
from oracle import model as M

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _plan(m, img, dtype, fused):
    """fused: 2 forces the fused kernel on every eligible grid (small test grids), 0 off."""
    return m.make_preconditioner(img, dtype, fused_mode=2 if fused else 0)


def _exact_model(n, C, mask, seed):
    """Block weights in {-2..2}/4 (C=1) or {-1..1}/4 (C=2, 3) and biases in {-2..2}/4, so
    every stencil value is a multiple of 1/16 below 16 in magnitude (C=3: at most
    81 * 3/16 + 1/2): exact in BF16/FP16 and in the split BF16 weights; c0 = 0.5, c1 = 1
    exactly."""
    from mlpcg_amd.sm_model_3d import SmallSMModelDn3D
    g = torch.Generator().manual_seed(seed)
    m = SmallSMModelDn3D(n, C, "trilinear", mask, False)
    span = 2 if C == 1 else 1
    with torch.no_grad():
        for k, p in m.named_parameters():
            if k.startswith(("c0", "c1")):
                p.zero_()
                if k.endswith("bias"):
                    p.fill_(0.5 if k.startswith("c0") else 1.0)
            elif k.endswith("weight"):
                p.copy_(torch.randint(-span, span + 1, p.shape, generator=g).float() / 4)
            else:
                p.copy_(torch.randint(-2, 3, p.shape, generator=g).float() / 4)
    return m.cuda()


def _exact_image(C, shape, seed, n):
    """Values in {0, 1/4, 1/2, 3/4} ({0, 1} for the mask channel), constant on 2^n cubes so
    that every pyramid level holds the same values and its stencils stay exact."""
    g = torch.Generator().manual_seed(seed)
    f = 2 ** n
    cs = tuple(d // f for d in shape)
    img = torch.randint(0, 4, (C, *cs), generator=g).float() / 4
    if C == 2:
        img[1] = (torch.rand(cs, generator=g) > 0.3).float()
    for ax in (1, 2, 3):
        img = img.repeat_interleave(f, dim=ax)
    return img.contiguous().cuda()


def _nrel(a, b):
    a, b = a.double(), b.double()
    return float((a - b).norm() / b.norm())


@pytest.mark.parametrize("C,mask,n,shape", [
    (1, False, 1, (16, 16, 128)),   # two fused levels, 2x2x2 and 1x1x1 tiles
    (1, False, 2, (24, 40, 192)),   # odd tile counts; level 1 (12, 20, 96) stays on stored K
    (2, True, 1, (16, 8, 64)),      # two channels, model mask
    (2, False, 2, (32, 16, 128)),
    (1, False, 1, (256, 128, 256)),  # 2048 tiles: 8-plane tiles on level 0, 4-plane on level 1
    (3, False, 2, (16, 16, 128)),   # three channels (the reference's image): 2-plane tiles
    (3, True, 1, (16, 8, 64)),      # three channels, model mask (fluid channel)
    (3, False, 1, (256, 128, 256)),  # three channels: tiles capped at 4 planes (LDS)
])
def test_fused_matches_stored_stencils(dev, C, mask, n, shape):
    """With exactly representable weights, images and input, both BF16 plans compute the
    same stencils exactly and differ only by FP32 summation order (plus the rare BF16
    rounding flip that causes): a wrong tap, halo, channel or lane mapping shows as O(1)."""
    dtype = torch.bfloat16
    m = _exact_model(n, C, mask, seed=3 + C)
    img = _exact_image(C, shape, seed=5, n=n)
    g = torch.Generator().manual_seed(7)
    r = torch.randint(-4, 5, shape, generator=g).float().cuda()
    pf, ps = _plan(m, img, dtype, True), _plan(m, img, dtype, False)
    assert pf.nbytes < ps.nbytes  # fused blocks keep no stencils
    # the level-0 pre block alone: exact sums on both paths, so bit-identical
    sc = torch.tensor(0.5, dtype=torch.float64, device=dev)
    yf, ys = pf.apply_block(0, r, dtype, sc), ps.apply_block(0, r, dtype, sc)
    assert yf.dtype == dtype and torch.equal(yf, ys)
    zf, zs = pf.apply(r), ps.apply(r)
    assert torch.isfinite(zf).all()
    assert _nrel(zf, zs) < 1e-5
    om = (img[0] > 0.3).to(torch.uint8)
    zf2, zs2 = pf.apply(r, in_scale=sc, out_mask=om), ps.apply(r, in_scale=sc, out_mask=om)
    assert _nrel(zf2, zs2) < 1e-5
    assert torch.all(zf2[om == 0] == 0)


@pytest.mark.parametrize("n,shape", [(2, (32, 16, 128)), (1, (64, 64, 128))])
def test_fused_fold_third_channel_exact(dev, n, shape):
    """Three-channel images [a, f, a + f] (the [air, fluid, non-solid] structure): plans
    whose fused blocks fold the third channel into the first two compute, with exactly
    representable weights and images, the same stencils as the three-channel plans, so
    every block and the whole application agree bit for bit."""
    m = _exact_model(n, 3, False, seed=21)
    g = torch.Generator().manual_seed(22)
    f = 2 ** n
    cs = tuple(d // f for d in shape)
    a = torch.randint(0, 3, cs, generator=g).float() / 4
    fl = torch.randint(0, 2, cs, generator=g).float() / 4
    img = torch.stack([a, fl, a + fl])
    for ax in (1, 2, 3):
        img = img.repeat_interleave(f, dim=ax)
    img = img.contiguous().cuda()
    r = torch.randint(-4, 5, shape, generator=torch.Generator().manual_seed(23)).float().cuda()
    p3 = m.make_preconditioner(img, torch.bfloat16, fused_mode=2)
    p2 = m.make_preconditioner(img, torch.bfloat16, fused_mode=2, fold_third=True)
    assert p2.fused_blocks == p3.fused_blocks > 0
    assert p2.nbytes < p3.nbytes  # two-channel FP16 image copies
    sc = torch.tensor(0.5, dtype=torch.float64, device=dev)
    assert torch.equal(p2.apply_block(0, r, torch.bfloat16, sc), p3.apply_block(0, r, torch.bfloat16, sc))
    assert torch.equal(p2.apply(r), p3.apply(r))


def test_fused_fold_third_channel_solve(dev):
    """npsd on the held-out 128^3 sphere scene with the shipped three-channel weights: the
    folded BF16 plan converges (FP64-checked residual) in the iterations of the
    three-channel BF16 plan, within one."""
    from mlpcg_amd.sm_model_3d import SmallSMModelDn3D
    from mlpcg_amd.solver import FusedPCG
    sd = torch.load(os.path.join(ROOT, "mlpcg_amd", "weights", "sm_l5_c3.pt"), weights_only=True)
    N = 128
    m = SmallSMModelDn3D(5, 3, "trilinear").cuda()
    m.load_state_dict(sd)
    sc = S.scene_spheres(N, 7)
    A = S.build_matrix(sc.flags)
    b = S.make_rhs(sc.flags, 0)[0]
    img = torch.as_tensor(S.image_channels(sc.flags, 3)).float().cuda()
    assert torch.equal(img[2], img[0] + img[1])
    s = FusedPCG(sc.flags, "flags", device=dev)
    bv = s.vector(b)
    its = {}
    for fold in (False, True):
        plan = m.make_preconditioner(img, torch.bfloat16, fold_third=fold)
        x, info = s.solve(bv, plan, method="npsd", max_it=300, tol=1e-6)
        assert info["status"] == 0, (fold, info)
        xs = x.cpu().numpy()[sc.fluid_cells].astype(np.float64)
        assert np.linalg.norm(b - A @ xs) / np.linalg.norm(b) < 1.05e-6
        its[fold] = info["iters"]
    assert abs(its[True] - its[False]) <= 1, its


@pytest.mark.parametrize("C,shape", [(1, (64, 64, 64)), (2, (64, 32, 128)), (3, (64, 64, 64))])
def test_block_forward_bf16_fused_exact(dev, C, shape):
    """The BF16 block forward (training and inference) on grids of 2^18 voxels or more runs
    the fused matrix-core apply.  With exactly representable weights, image and input every
    stencil and every sum is exact on both paths, so y equals the FP64 forward rounded to
    BF16 bit for bit: a wrong tap, halo, channel or weight packing shows at once."""
    from mlpcg_amd import kernels as K
    g = torch.Generator().manual_seed(11 + C)
    span = 2 if C == 1 else 1
    W = torch.randint(-span, span + 1, (27, C, 3, 3, 3), generator=g).double() / 4
    b = torch.randint(-2, 3, (27,), generator=g).double() / 4
    img = torch.randint(0, 4, (C, *shape), generator=g).double() / 4
    x = torch.randint(-4, 5, (1, 1, *shape), generator=g).double()
    y64 = K.sm_block_forward(img.to(dev), x.to(dev), W.to(dev), b.to(dev))
    y16 = K.sm_block_forward(img.to(dev).bfloat16(), x.to(dev).bfloat16(), W.float().to(dev),
                             b.float().to(dev))
    assert y16.dtype == torch.bfloat16
    assert torch.equal(y16, y64.bfloat16())


def test_fused_plan_vs_fp64_trained(dev):
    """Trained weights on a 64^3 pillars scene: the fused BF16 plan is about as close to
    the FP64 oracle composition as the stored-K BF16 plan (both round the stencils to
    BF16; the fused plan also rounds its level-0 input)."""
    dtype, tol = torch.bfloat16, 6e-3
    from mlpcg_amd.sm_model_3d import SmallSMModelDn3D
    sd = torch.load(os.path.join(ROOT, "mlpcg_amd", "weights", "sm_l5_c1.pt"), weights_only=True)
    n, N = 5, 64
    m = SmallSMModelDn3D(n, 1, "trilinear").cuda()
    m.load_state_dict(sd)
    sc = S.scene_pillars(N, 0)
    img = torch.as_tensor(S.image_channels(sc.flags, 1))
    rng = np.random.default_rng(2)
    r = torch.as_tensor(rng.standard_normal(N ** 3) * (sc.flags == S.FLUID).ravel()).reshape(N, N, N)
    P64 = {k: v.double() for k, v in sd.items()}
    ref = M.model_forward(P64, n, img.double(), r.double().reshape(1, 1, N, N, N)).reshape(N, N, N)
    pf = _plan(m, img.float().cuda(), dtype, True)
    ps = _plan(m, img.float().cuda(), dtype, False)
    ef = _nrel(pf.apply(r.float().cuda()).cpu(), ref)
    es = _nrel(ps.apply(r.float().cuda()).cpu(), ref)
    assert ef < tol, (ef, es)
    assert ef < 2.5 * es + 1e-4, (ef, es)


@pytest.mark.parametrize("C", [1, 3])
def test_fused_bf16_solve_iterations(dev, C):
    """npsd on the held-out 128^3 sphere scene of the shipped weights (diag(A)/6 image, or
    the three binary channels [air, fluid, non-solid]): the fused BF16 plan converges
    (FP64-checked residual) within 1.1x the iterations of the FP32 plan, and the stored-K
    BF16 plan within 1.1x too."""
    from mlpcg_amd.sm_model_3d import SmallSMModelDn3D
    from mlpcg_amd.solver import FusedPCG
    sd = torch.load(os.path.join(ROOT, "mlpcg_amd", "weights", f"sm_l5_c{C}.pt"), weights_only=True)
    N = 128
    m = SmallSMModelDn3D(5, C, "trilinear").cuda()
    m.load_state_dict(sd)
    sc = S.scene_spheres(N, 7)
    A = S.build_matrix(sc.flags)
    b = S.make_rhs(sc.flags, 0)[0]
    img = torch.as_tensor(S.image_channels(sc.flags, C)).float().cuda()
    s = FusedPCG(sc.flags, "flags", device=dev)
    bv = s.vector(b)
    its = {}
    for name, dt, fused in (("fp32", torch.float32, True), ("bf16", torch.bfloat16, True),
                            ("bf16_stored", torch.bfloat16, False)):
        plan = _plan(m, img, dt, fused)
        assert (plan.fused_blocks > 0) == (name == "bf16")
        x, info = s.solve(bv, plan, method="npsd", max_it=300, tol=1e-6)
        assert info["status"] == 0, (name, info)
        xs = x.cpu().numpy()[sc.fluid_cells].astype(np.float64)
        assert np.linalg.norm(b - A @ xs) / np.linalg.norm(b) < 1.05e-6
        its[name] = info["iters"]
    assert its["bf16"] <= 1.1 * its["fp32"] + 1, its
    assert its["bf16_stored"] <= 1.1 * its["fp32"] + 1, its


@pytest.mark.parametrize("method", ["npcg", "npcg_flex", "npsd"])
def test_fused_epilogue_dots_in_solver(dev, method):
    """With a fused last block the solver's dot products on z (r.z, r_old.z; q.AN, q.AO)
    come from the apply's epilogue instead of separate kernels.  With exactly representable
    stencils both plans give the same z up to FP32 summation order, so the residual
    histories must agree; a wrong pairing of dots would separate them at once."""
    from mlpcg_amd.solver import FusedPCG
    N = 64
    sc = S.scene_spheres(N, 11, num_spheres=3, rmin=4, rmax=9)
    b = S.make_rhs(sc.flags, 2)[0]
    m = _exact_model(2, 1, False, seed=9)
    img = _exact_image(1, (N, N, N), seed=4, n=2)
    s = FusedPCG(sc.flags, "flags", device=dev)
    bv = s.vector(b)
    hist = {}
    for fused in (True, False):
        plan = _plan(m, img, torch.bfloat16, fused)
        _, info = s.solve(bv, plan, method=method, max_it=12, tol=1e-12, history=True)
        assert info["iters"] == 12
        hist[fused] = info["history"]
    rel = np.abs(hist[True] - hist[False]) / hist[False]
    assert rel.max() < 1e-3, rel
