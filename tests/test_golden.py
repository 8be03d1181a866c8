"""Golden fixtures (tests/golden/*.npz): the oracle reproduces them (CPU), and the HIP
path matches them through the C ABI (GPU)."""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import torch

from mlpcg_amd import scenes as S
from oracle import model as M
from oracle import solvers as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def _t(a):
    return torch.as_tensor(np.asarray(a))


def _params(z, prefix):
    p = f"{prefix}_p_"
    return {k[len(p):]: _t(z[k]) for k in z.files if k.startswith(p)}


MODEL_CASES = [("m_c3_mask", 2, "trilinear", True, False), ("m_c1_swap", 3, "trilinear", False, True),
               ("m_c3_nearest", 2, "nearest", False, False)]


# ----------------------------------------------------------------------------- CPU

@pytest.mark.parametrize("C", [1, 3])
def test_oracle_blocks_reproduce_golden(C):
    z = _load("sm_blocks.npz")
    img, x, W, b = (_t(z[f"c{C}_{k}"]) for k in ("img", "x", "W", "b"))
    assert np.allclose(M.sm_block(img, x, W, b).numpy(), z[f"c{C}_block_y"], rtol=0, atol=1e-13)
    assert np.allclose(M.sm_block_trans(img, x, W, b).numpy(), z[f"c{C}_trans_y"], rtol=0, atol=1e-13)
    zz, yb = M.sm_linear(img, _t(z[f"c{C}_lin_w"]), _t(z[f"c{C}_lin_b"]))
    assert np.allclose(yb.numpy(), z[f"c{C}_lin_ybar"], atol=1e-14)


@pytest.mark.parametrize("name,n,mode,mask,swap", MODEL_CASES)
def test_oracle_model_reproduces_golden(name, n, mode, mask, swap):
    z = _load("sm_models.npz")
    y = M.model_forward(_params(z, name), n, _t(z[f"{name}_img"]), _t(z[f"{name}_b"]), mode, mask, swap)
    assert np.allclose(y.numpy(), z[f"{name}_y"], rtol=0, atol=1e-12)


def test_matrix_and_solvers_reproduce_golden():
    z = _load("laplacian_pcg.npz")
    A = S.build_matrix(z["flags"])
    G = sp.csr_matrix((z["data"], z["indices"], z["indptr"]), shape=A.shape)
    assert abs(A - G).max() == 0
    d = A.diagonal()
    for name, f in (("npcg", O.npcg), ("npcg_flex", O.npcg_flex), ("npsd", O.npsd)):
        x, k, hist = f(z["b"], A, np.zeros_like(z["b"]), lambda r: r / d, 400, tol=1e-6)
        assert k == int(z[f"{name}_k"])
        assert np.allclose(hist, z[f"{name}_hist"], rtol=1e-10)


# ----------------------------------------------------------------------------- GPU

@pytest.mark.gpu
@pytest.mark.parametrize("C", [1, 3])
@pytest.mark.parametrize("trans", [False, True])
def test_hip_blocks_match_golden(dev, C, trans):
    from mlpcg_amd import kernels as K
    z = _load("sm_blocks.npz")
    img, x, W, b, g = (_t(z[f"c{C}_{k}"]).to(dev) for k in ("img", "x", "W", "b", "g"))
    tag = f"c{C}_{'trans' if trans else 'block'}"
    f = K.sm_block_trans_forward if trans else K.sm_block_forward
    bw = K.sm_block_trans_backward if trans else K.sm_block_backward
    y = f(img, x, W, b).cpu().numpy()
    assert np.allclose(y, z[f"{tag}_y"], rtol=0, atol=1e-13)
    gx, gW, gb = bw(g, img, x, W, b)
    for got, key in ((gx, "gx"), (gW, "gW"), (gb, "gb")):
        ref = z[f"{tag}_{key}"]
        assert np.allclose(got.cpu().numpy(), ref, rtol=0, atol=1e-12 * max(1.0, np.abs(ref).max()))
    zz, yb = K.sm_linear(img, _t(z[f"c{C}_lin_w"]).to(dev), _t(z[f"c{C}_lin_b"]).to(dev))
    assert np.allclose(yb.cpu().numpy(), z[f"c{C}_lin_ybar"], atol=1e-14)
    assert np.allclose(zz.cpu().numpy(), z[f"c{C}_lin_z"], atol=1e-14)


@pytest.mark.gpu
@pytest.mark.parametrize("name,n,mode,mask,swap", MODEL_CASES)
def test_hip_model_and_plan_match_golden(dev, name, n, mode, mask, swap):
    from mlpcg_amd.sm_model_3d import SmallSMModelDn3D
    z = _load("sm_models.npz")
    P = _params(z, name)
    C = int(z[f"{name}_img"].shape[0])
    m = SmallSMModelDn3D(n, C, mode, mask, swap).to(dev, torch.float64)
    with torch.no_grad():
        for k, p in m.named_parameters():
            p.copy_(P[k])
        y = m(_t(z[f"{name}_img"]).to(dev), _t(z[f"{name}_b"]).to(dev)).cpu().numpy()
    assert np.allclose(y, z[f"{name}_y"], rtol=0, atol=1e-11)
    plan = m.make_preconditioner(_t(z[f"{name}_img"]).to(dev), torch.float64)
    zp = plan.apply(_t(z[f"{name}_b"]).to(dev).reshape(z[f"{name}_b"].shape[-3:])).cpu().numpy()
    assert np.allclose(zp.reshape(y.shape), z[f"{name}_y"], rtol=0, atol=1e-11)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["spd", "smallspd"])
def test_hip_spd_plan_matches_golden(dev, name):
    from mlpcg_amd.sm_model_3d import SPDSMModelDn3D, SmallSPDSMModelDn3D
    z = _load("sm_models.npz")
    P = _params(z, name)
    m = (SPDSMModelDn3D if name == "spd" else SmallSPDSMModelDn3D)(2, 3).to(dev, torch.float64)
    with torch.no_grad():
        for k, p in m.named_parameters():
            p.copy_(P[k])
    img, b = _t(z[f"{name}_img"]).to(dev), _t(z[f"{name}_b"]).to(dev)
    zp = m.make_preconditioner(img, torch.float64).apply(b.reshape(b.shape[-3:])).cpu().numpy()
    assert np.allclose(zp.reshape(z[f"{name}_y"].shape), z[f"{name}_y"], rtol=0, atol=1e-11)


@pytest.mark.gpu
def test_hip_spmv_matches_golden_matrix(dev):
    from mlpcg_amd import kernels as K
    from mlpcg_amd.solver import CSRMatrix
    z = _load("laplacian_pcg.npz")
    A = sp.csr_matrix((z["data"], z["indices"], z["indptr"]))
    x = np.arange(A.shape[0], dtype=np.float32) % 17 - 8
    ref = (A @ x.astype(np.float64)).astype(np.float32)
    y = (CSRMatrix(A, dev) @ torch.as_tensor(x, device=dev)).cpu().numpy()
    assert np.array_equal(y, ref)
    flags = z["flags"]
    code = K.flags_code(torch.as_tensor(flags, device=dev))
    xg = torch.as_tensor(S.expand(x, flags).ravel(), device=dev)
    yf = K.spmv_flags(code, xg).cpu().numpy()[flags.ravel() == S.FLUID]
    assert np.array_equal(yf, ref)


# --------------------------------------------- fixtures written by the reference's own model

REF_CASES = ["c1_n2_8", "c1_n3_16", "c3_n2_16", "c3_n3_16"]


@pytest.mark.parametrize("name", REF_CASES)
def test_oracle_model_matches_reference_model(name):
    """The oracle's composition (nearest upsample, no mask, no swap) reproduces
    SmallSMModelDn3DPY (sm_model_3d_py.py:33-63) run by tests/golden/make_ref_model.py,
    and its SMLinear coefficients equal the reference's Conv3d(C->27).mean() per level."""
    z = _load("ref_models.npz")
    n = int(z[f"{name}_n"])
    P = _params(z, name)
    img, b = _t(z[f"{name}_img"]), _t(z[f"{name}_b"])
    y = M.model_forward(P, n, img, b, mode="nearest", mask=False, swap=False)
    ref = z[f"{name}_y"]
    assert np.abs(y.numpy() - ref).max() <= 1e-12 * np.abs(ref).max()
    im = img
    for i in range(n):
        for which in ("c0", "c1"):
            c, _ = M.sm_linear(im, P[f"{which}.{i}.weight"], P[f"{which}.{i}.bias"])
            assert abs(float(c) - float(z[f"{name}_{which}_{i}"])) < 1e-13
        im = M.avg_pool(im)


@pytest.mark.gpu
@pytest.mark.parametrize("name", REF_CASES)
def test_hip_model_and_plan_match_reference_model(dev, name):
    """The drop-in module (batched forward) and the native plan (one vector at a time)
    in FP64 reproduce the reference model's outputs."""
    from mlpcg_amd.sm_model_3d import SmallSMModelDn3D
    z = _load("ref_models.npz")
    n = int(z[f"{name}_n"])
    P = _params(z, name)
    img, b, ref = _t(z[f"{name}_img"]).to(dev), _t(z[f"{name}_b"]).to(dev), z[f"{name}_y"]
    C = int(img.shape[0])
    m = SmallSMModelDn3D(n, C, "nearest", False, False).to(dev, torch.float64)
    with torch.no_grad():
        for k, p in m.named_parameters():
            p.copy_(P[k])
        y = m(img, b).cpu().numpy()
    scale = np.abs(ref).max()
    assert np.abs(y - ref).max() <= 1e-11 * scale
    plan = m.make_preconditioner(img, torch.float64)
    for bi in range(b.shape[0]):
        zp = plan.apply(b[bi, 0].contiguous()).cpu().numpy()
        assert np.abs(zp - ref[bi, 0]).max() <= 1e-11 * scale
