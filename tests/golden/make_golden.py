"""Write the small golden fixtures of this repository (tests/golden/*.npz).

The reference ships no golden vectors, no checkpoints and no test data for this path
(its tests/ and data directories are git-ignored), so these fixtures are produced by
the FP64 oracle of this repository, after the oracle itself is pinned by the
brute-force, library-routine and closed-form checks of tests/test_oracle_*.py.  They
freeze those results so the oracle and the HIP path are both checked against fixed
numbers.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from mlpcg_amd import scenes as S  # noqa: E402
from oracle import model as M  # noqa: E402
from oracle import solvers as O  # noqa: E402


def _rand(shape, seed):
    g = torch.Generator().manual_seed(seed)
    return torch.rand(*shape, generator=g, dtype=torch.float64) * 2 - 1


def blocks():
    out = {}
    for C in (1, 3):
        shape = (6, 5, 8)
        img, x, g = _rand((C, *shape), 10 + C), _rand((2, 1, *shape), 20 + C), _rand((2, 1, *shape), 30 + C)
        W, b = _rand((27, C, 3, 3, 3), 40 + C) / 9, _rand((27,), 50 + C) / 9
        out[f"c{C}_img"], out[f"c{C}_x"], out[f"c{C}_g"] = img.numpy(), x.numpy(), g.numpy()
        out[f"c{C}_W"], out[f"c{C}_b"] = W.numpy(), b.numpy()
        for trans in (False, True):
            tag = f"c{C}_{'trans' if trans else 'block'}"
            xr, Wr, br = x.clone().requires_grad_(), W.clone().requires_grad_(), b.clone().requires_grad_()
            y = (M.sm_block_trans if trans else M.sm_block)(img, xr, Wr, br)
            (y * g).sum().backward()
            out[f"{tag}_y"] = y.detach().numpy()
            out[f"{tag}_gx"], out[f"{tag}_gW"], out[f"{tag}_gb"] = (
                xr.grad.numpy(), Wr.grad.numpy(), br.grad.numpy())
        w1, b1 = _rand((1, C, 3, 3, 3), 60 + C), _rand((1,), 70 + C)
        z, ybar = M.sm_linear(img, w1, b1)
        out[f"c{C}_lin_w"], out[f"c{C}_lin_b"] = w1.numpy(), b1.numpy()
        out[f"c{C}_lin_z"], out[f"c{C}_lin_ybar"] = z.numpy(), ybar.numpy()
    np.savez_compressed(os.path.join(HERE, "sm_blocks.npz"), **out)


def model():
    out = {}
    cases = [("m_c3_mask", 2, 3, (8, 8, 8), "trilinear", True, False),
             ("m_c1_swap", 3, 1, (16, 8, 16), "trilinear", False, True),
             ("m_c3_nearest", 2, 3, (8, 16, 8), "nearest", False, False)]
    for name, n, C, shape, mode, mask, swap in cases:
        P = M.init_params(n, C, torch.Generator().manual_seed(n * 10 + C))
        img = _rand((C, *shape), 80).abs() if C == 1 else (_rand((C, *shape), 81) > -0.2).double()
        b = _rand((1, 1, *shape), 82)
        y = M.model_forward(P, n, img, b, mode, mask, swap)
        out[f"{name}_img"], out[f"{name}_b"], out[f"{name}_y"] = img.numpy(), b.numpy(), y.numpy()
        for k, v in P.items():
            out[f"{name}_p_{k}"] = v.numpy()
    for name, reduced in (("spd", False), ("smallspd", True)):
        n, C, shape = 2, 3, (8, 8, 16)
        P = M.init_params_spd(n, C, torch.Generator().manual_seed(90 + int(reduced)), reduced=reduced)
        img = (_rand((C, *shape), 91) > -0.2).double()
        b = _rand((1, 1, *shape), 92)
        y = (M.small_spd_forward if reduced else M.spd_forward)(P, n, img, b)
        out[f"{name}_img"], out[f"{name}_b"], out[f"{name}_y"] = img.numpy(), b.numpy(), y.numpy()
        for k, v in P.items():
            out[f"{name}_p_{k}"] = v.numpy()
    np.savez_compressed(os.path.join(HERE, "sm_models.npz"), **out)


def laplacian_and_solvers():
    sc = S.scene_spheres(16, 3, num_spheres=2, rmin=2, rmax=4)
    A = S.build_matrix(sc.flags)
    b = S.make_rhs(sc.flags, 3)[0]
    d = A.diagonal()
    jac = lambda r: r / d
    out = {"flags": sc.flags, "indptr": A.indptr, "indices": A.indices, "data": A.data, "b": b}
    for name, f in (("npcg", O.npcg), ("npcg_flex", O.npcg_flex), ("npsd", O.npsd)):
        x, k, hist = f(b, A, np.zeros_like(b), jac, 400, tol=1e-6)
        out[f"{name}_x"], out[f"{name}_k"], out[f"{name}_hist"] = x, np.array(k), np.array(hist)
    np.savez_compressed(os.path.join(HERE, "laplacian_pcg.npz"), **out)


if __name__ == "__main__":
    blocks()
    model()
    laplacian_and_solvers()
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))
