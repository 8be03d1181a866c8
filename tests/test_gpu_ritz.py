"""GPU parity of the Ritz-vector pipeline (mlpcg_amd/ritz.py) with the FP64 restatement of
lib/create_dataset.py:28-70 and preprocess.py:38-58 (oracle/ritz.py), and the
mathematical properties the oracle is pinned by."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla
import torch

from mlpcg_amd import scenes as S
from oracle import ritz as R

pytestmark = pytest.mark.gpu


def _problem(n=16, seed=3):
    sc = S.scene_spheres(n, seed, num_spheres=1, rmin=2, rmax=3)
    A = S.build_matrix(sc.flags)
    b = S.make_rhs(sc.flags, seed)[0]
    return sc, A, b


def test_fp64_kernels(dev):
    from mlpcg_amd import ritz as G
    n = 1001  # not a multiple of the 8-double run: exercises the tail
    g = torch.Generator().manual_seed(0)
    a = torch.randn(n, generator=g, dtype=torch.float64)
    b = torch.randn(n, generator=g, dtype=torch.float64)
    ops = G._Vec(dev)
    out = torch.zeros(4, dtype=torch.float64, device=dev)
    ad, bd = a.to(dev), b.to(dev)
    ops.dot(ad, bd, out[0:1])
    assert abs(float(out[0]) - float(a @ b)) < 1e-12 * float(a.abs() @ b.abs())
    out[1] = 0.5
    out[2] = -0.25
    w = ad.clone()
    ops.update(w, bd, ad, out[1:2], out[2:3])
    assert torch.allclose(w.cpu(), a - 0.5 * b + 0.25 * a, rtol=0, atol=1e-15)
    ops.dot(w, w, out[3:4])
    v = torch.empty_like(w)
    nrm = torch.zeros(1, dtype=torch.float64, device=dev)
    ops.normalize(v, w, out[3:4], nrm)
    assert abs(float(v.norm()) - 1.0) < 1e-14 and abs(float(nrm) - float(w.norm())) < 1e-12


def test_lanczos_matches_oracle(dev):
    from mlpcg_amd import ritz as G
    _, A, b = _problem()
    m = 60
    V_ref, a_ref, e_ref = R.lanczos(A, b, m)
    Ad = G.CSRMatrix64(A, dev)
    V, a, e = G.lanczos(Ad, b, m)
    assert V.shape == V_ref.shape
    # same recurrence, FP64 both sides; summation orders differ, so rounding differences
    # grow slowly with the step count
    assert np.allclose(a, a_ref, rtol=0, atol=1e-8 * np.abs(a_ref).max())
    assert np.allclose(e, e_ref, rtol=0, atol=1e-8 * np.abs(e_ref).max())
    Vh = V.cpu().numpy()
    assert np.abs(Vh @ Vh.T - np.eye(m)).max() < 1e-12
    assert np.abs(Vh - V_ref).max() < 1e-6


def test_ritz_vectors_and_rhs_match_oracle(dev):
    from mlpcg_amd import ritz as G
    sc, A, b = _problem(n=16, seed=4)
    m = 80
    vals_ref, vecs_ref = R.create_ritz_vec(A, b, m)
    Ad = G.CSRMatrix64(A, dev)
    vals, Rv = G.ritz_vectors(Ad, b, m)
    assert np.allclose(vals, vals_ref, rtol=0, atol=1e-9 * np.abs(vals_ref).max())
    # the largest Ritz values are eigenvalues of A
    top = spla.eigsh(A, k=2, which="LA", return_eigenvectors=False)
    assert np.allclose(np.sort(vals)[-2:], np.sort(top), rtol=1e-8)
    Rh = Rv.cpu().numpy()
    assert np.abs(Rh @ Rh.T - np.eye(len(vals))).max() < 1e-10
    # well-separated extreme vectors agree up to sign
    for i in (0, len(vals) - 1):
        s = np.sign(Rh[i] @ vecs_ref[i])
        assert np.abs(s * Rh[i] - vecs_ref[i]).max() < 1e-6
    coef = np.random.default_rng(1).standard_normal((len(vals), 6))
    B_ref = R.training_rhs(vecs_ref, coef, num_requested=m)
    B = G.training_rhs(Rv, coef=coef, num_requested=m, dtype=torch.float64).cpu().numpy()
    # Ritz vectors are defined up to sign, the coefficients each right-hand side carries in
    # its own Ritz basis are not: both recover the recipe's weighted coefficients
    rec = Rh @ B.T
    rec_ref = vecs_ref @ B_ref.T
    assert np.abs(rec - rec_ref).max() < 1e-6
    assert np.allclose(np.linalg.norm(B, axis=1), 1.0)
    # drawn on the device: unit norm, inside the Ritz span
    Bd = G.training_rhs(Rv, count=4, num_requested=m, generator=torch.Generator(device=dev).manual_seed(0))
    Bh = Bd.double().cpu().numpy()
    assert np.allclose(np.linalg.norm(Bh, axis=1), 1.0, atol=1e-6)
    assert np.abs(Bh - (Rh.T @ (Rh @ Bh.T)).T).max() < 1e-5
    full = G.full_grid(Bd, torch.as_tensor(sc.fluid_cells, device=dev), sc.flags.size)
    assert full.shape == (4, sc.flags.size)
    assert float(full.abs().sum()) == pytest.approx(float(Bd.abs().sum()), rel=1e-6)
