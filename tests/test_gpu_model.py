"""GPU parity of the drop-in sm_model_3d modules and of the native plan against the FP64
oracle composition (oracle/model.py)."""
import pytest
import torch

from oracle import model as M

pytestmark = pytest.mark.gpu


def _rand(shape, seed, dtype=torch.float64):
    g = torch.Generator().manual_seed(seed)
    return (torch.rand(*shape, generator=g, dtype=torch.float64) * 2 - 1).to(dtype)


def _rel(a, b):
    a = a.detach().double().cpu()
    b = b.detach().double().cpu()
    return float((a - b).abs().max() / max(b.abs().max(), 1e-300))


def _image(C, shape, seed):
    if C == 1:
        return _rand((1, *shape), seed).abs()
    return (_rand((C, *shape), seed) > -0.2).double()


def _small_model(n, C, mode="trilinear", mask=False, swap=False, seed=0, dtype=torch.float64, dev="cuda"):
    from mlpcg_amd.sm_model_3d import SmallSMModelDn3D
    P = M.init_params(n, C, torch.Generator().manual_seed(seed))
    m = SmallSMModelDn3D(n, C, mode, mask, swap)
    m.load_state_dict({k: v.to(torch.float32) for k, v in P.items()})
    m = m.to(dev, dtype)
    with torch.no_grad():  # exact FP64 copy of the oracle parameters
        for k, p in m.named_parameters():
            p.copy_(P[k].to(dtype))
    return m, P


@pytest.mark.parametrize("n,C,shape,mode,mask,swap", [
    (2, 1, (8, 8, 8), "trilinear", False, False),
    (3, 3, (16, 8, 16), "trilinear", True, False),
    (2, 3, (8, 16, 8), "nearest", False, True),
    (3, 1, (16, 16, 16), "trilinear", False, True),
])
def test_model_forward_f64(dev, n, C, shape, mode, mask, swap):
    m, P = _small_model(n, C, mode, mask, swap, seed=n + C)
    img, b = _image(C, shape, 1), _rand((2, 1, *shape), 2)
    ref = M.model_forward(P, n, img, b, mode, mask, swap)
    with torch.no_grad():
        y = m(img.to(dev), b.to(dev))
        ye = m.eval_forward(img.to(dev), b.to(dev), None, [], [], [])
    assert _rel(y, ref) < 1e-11
    assert _rel(ye, ref) < 1e-11


@pytest.mark.parametrize("n,C,shape", [(4, 1, (64, 64, 64)), (3, 3, (32, 32, 32))])
def test_model_forward_f32_tolerance(dev, n, C, shape):
    """FP32 preconditioner output within 1e-5 relative of the FP64 oracle."""
    m, P = _small_model(n, C, seed=7, dtype=torch.float32)
    img, b = _image(C, shape, 3), _rand((1, 1, *shape), 4)
    Pf = {k: v.float().double() for k, v in P.items()}
    ref = M.model_forward(Pf, n, img.float().double(), b.float().double())
    with torch.no_grad():
        y = m.eval_forward(img.float().to(dev), b.float().to(dev), None, [], [], [])
    assert _rel(y, ref) < 1e-5


@pytest.mark.parametrize("kind", ["spd", "small_spd"])
def test_spd_models_f64(dev, kind):
    from mlpcg_amd.sm_model_3d import SPDSMModelDn3D, SmallSPDSMModelDn3D
    n, C, shape = 2, 3, (8, 8, 16)
    reduced = kind == "small_spd"
    P = M.init_params_spd(n, C, torch.Generator().manual_seed(5), reduced=reduced)
    m = (SmallSPDSMModelDn3D if reduced else SPDSMModelDn3D)(n, C).to(dev, torch.float64)
    with torch.no_grad():
        for k, p in m.named_parameters():
            p.copy_(P[k])
    img, b = _image(C, shape, 6), _rand((1, 1, *shape), 7)
    f = M.small_spd_forward if reduced else M.spd_forward
    ref = f(P, n, img, b)
    with torch.no_grad():
        y = m(img.to(dev), b.to(dev))
    assert _rel(y, ref) < 1e-11
    # native plan of the SPD model, FP64 storage
    plan = m.make_preconditioner(img.to(dev), torch.float64)
    z = plan.apply(b.to(dev).reshape(shape))
    assert _rel(z.reshape(b.shape), ref) < 1e-11


def test_model_backward_f64(dev):
    """Parameter and input gradients of the whole model vs autograd of the oracle."""
    n, C, shape = 2, 3, (8, 8, 8)
    m, P = _small_model(n, C, "trilinear", True, False, seed=11)
    img, b, R = _image(C, shape, 8), _rand((2, 1, *shape), 9), _rand((2, 1, *shape), 10)
    bb = b.clone().to(dev).requires_grad_()
    (m(img.to(dev), bb) * R.to(dev)).sum().backward()
    Pr = {k: v.clone().requires_grad_() for k, v in P.items()}
    br = b.clone().requires_grad_()
    (M.model_forward(Pr, n, img, br, "trilinear", True, False) * R).sum().backward()
    assert _rel(bb.grad, br.grad) < 1e-10
    for k, p in m.named_parameters():
        assert _rel(p.grad, Pr[k].grad) < 1e-10, k


@pytest.mark.parametrize("dtype,tol", [(torch.float32, 1e-5), (torch.bfloat16, 3e-2), (torch.float16, 5e-3)])
def test_plan_apply(dev, dtype, tol):
    """Native plan (K cached, fused epilogues) vs the FP64 oracle composition."""
    n, C, shape = 3, 1, (32, 32, 32)
    m, P = _small_model(n, C, seed=12, dtype=torch.float32)
    img, r = _image(C, shape, 13), _rand((1, 1, *shape), 14).float()
    Pf = {k: v.float().double() for k, v in P.items()}
    ref = M.model_forward(Pf, n, img.float().double(), r.double())
    plan = m.make_preconditioner(img.float().to(dev), dtype)
    z = plan.apply(r.to(dev).reshape(shape))
    assert z.dtype == torch.float32
    assert _rel(z.reshape(ref.shape), ref) < tol
    # input scale and output mask
    sc = torch.tensor(3.0, dtype=torch.float64, device=dev)
    mask = (img[0] > 0.5).to(torch.uint8).to(dev)
    z2 = plan.apply(r.to(dev).reshape(shape), in_scale=sc, out_mask=mask)
    ref2 = torch.where(mask.cpu().bool(), 3.0 * ref.reshape(shape), torch.zeros(()))
    assert _rel(z2, ref2) < tol


def test_plan_matches_module_eval_forward_f32(dev):
    n, C, shape = 4, 3, (32, 32, 32)
    m, _ = _small_model(n, C, "trilinear", True, False, seed=15, dtype=torch.float32)
    img, r = _image(C, shape, 16).float(), _rand((1, 1, *shape), 17).float()
    with torch.no_grad():
        ye = m.eval_forward(img.to(dev), r.to(dev), None, [], [], [])
    z = m.make_preconditioner(img.to(dev), torch.float32).apply(r.to(dev).reshape(shape))
    assert _rel(z.reshape(ye.shape), ye) < 1e-6


@pytest.mark.parametrize("dtype,tol", [(torch.float32, 2e-6), (torch.bfloat16, 2e-3), (torch.float16, 5e-4)])
@pytest.mark.parametrize("n,shape,mode", [(3, (32, 32, 32), "trilinear"), (2, (16, 32, 16), "nearest"),
                                          (4, (64, 64, 64), "trilinear"), (2, (16, 16, 12), "trilinear")])
def test_coarse_tail_matches_kernel_sequence(dev, dtype, tol, n, shape, mode):
    """The one-workgroup kernel for the two coarsest levels computes the same V-cycle tail
    as the seven per-level kernels (same roundings to the storage type; FP32 summation
    order may differ).  Level n-1 of (16, 16, 12) has n3 = 6: runs of 2 z cells."""
    m, _ = _small_model(n, 1, mode, seed=31, dtype=torch.float32)
    img, r = _image(1, shape, 32).float().to(dev), _rand((1, 1, *shape), 33).float().to(dev)
    z = {}
    for tail in (True, False):
        plan = m.make_preconditioner(img, dtype, coarse_tail=tail)
        assert plan.coarse_tail == tail
        z[tail] = plan.apply(r.reshape(shape)).double()
    assert float((z[True] - z[False]).norm() / z[False].norm()) < tol


@pytest.mark.parametrize("act", [torch.float32, torch.bfloat16])
def test_forward_cached_matches_forward(dev, act):
    """forward_cached (image pyramid and SMLinear mean patches computed once per scene)
    gives forward's output and parameter gradients."""
    n, C, shape = 3, 1, (32, 32, 32)
    m, _ = _small_model(n, C, seed=41, dtype=torch.float32)
    img = _image(C, shape, 42).float().to(dev)
    b = _rand((2, 1, *shape), 43).float().to(dev).to(act)
    R = _rand((2, 1, *shape), 44).float().to(dev)
    outs, grads = [], []
    for cached in (False, True):
        m.zero_grad(set_to_none=True)
        y = m.forward_cached(m.image_cache(img, act), b) if cached else m(img.to(act), b)
        (y.float() * R).sum().backward()
        outs.append(y.detach().float())
        grads.append({k: p.grad.detach().clone() for k, p in m.named_parameters()})
    tol = 1e-6 if act == torch.float32 else 1e-2
    assert _rel(outs[1], outs[0]) < tol
    for k in grads[0]:
        assert _rel(grads[1][k], grads[0][k]) < (1e-5 if act == torch.float32 else 5e-2), k
