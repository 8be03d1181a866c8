"""Write tests/golden/ref_models.npz from the reference's own pure-PyTorch model.

The composition of SmallSMModelDn3D (pool / pre / l / upsample / post / level combine) is
the one part of the oracle that no library routine, closed form or brute-force check
pins independently, so its fixtures come from the reference itself:
SmallSMModelDn3DPY (sm_model_3d_py.py:33-63, nearest-neighbour upsample, no mask, no
swap) is imported from the reference tree IN THE BUILD CONTAINER ONLY, run in float64 on
the CPU, and its inputs, weights and outputs are saved as plain arrays.  Nothing of the
reference is stored: only seeded inputs, weights and the outputs it computed.

Weights are given to the fixture in this repository's state_dict layout:
  pre.i / post.i / l  : KL.weight (27, C, 3, 3, 3), KL.bias (27,)   (SmallSMBlock3DPY)
  c0.i / c1.i         : KL.weight.mean(0, keepdim=True) (1, C, 3, 3, 3),
                        KL.bias.mean(0, keepdim=True) (1,)
The second mapping is exact: SmallLinearBlock3DPY returns Conv3d(C -> 27)(image).mean(),
and the mean over the 27 output channels of a linear map is the map with the averaged
weight and bias (sm_model_3d_py.py:24-31 against sm_model_3d.py:147-157).

    python tests/golden/make_ref_model.py [/path/to/reference]
"""
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"

# (name, levels n, channels C, grid)
CASES = [
    ("c1_n2_8", 2, 1, (8, 8, 8)),
    ("c1_n3_16", 3, 1, (16, 8, 16)),
    ("c3_n2_16", 2, 3, (16, 8, 8)),
    ("c3_n3_16", 3, 3, (8, 16, 16)),
]


def main():
    sys.path.insert(0, REF)
    from sm_model_3d_py import SmallSMModelDn3DPY  # noqa: E402  (reference, read-only)

    out = {}
    for ci, (name, n, C, shape) in enumerate(CASES):
        torch.manual_seed(1000 + ci)
        model = SmallSMModelDn3DPY(n, C).double().eval()
        g = torch.Generator().manual_seed(2000 + ci)
        if C == 1:
            img = torch.rand((1, 1, *shape), generator=g, dtype=torch.float64)
        else:  # binary [air, fluid, non-solid] channels like lib/read_data.py:28-39
            u = torch.rand(shape, generator=g, dtype=torch.float64)
            air, fluid = (u > 0.8).double(), ((u > 0.25) & (u <= 0.8)).double()
            img = torch.stack([air, fluid, air + fluid])[None]
        b = torch.randn((2, 1, *shape), generator=g, dtype=torch.float64)
        with torch.no_grad():
            y = model(img[0], b)
        out[f"{name}_img"] = img[0].numpy()
        out[f"{name}_b"] = b.numpy()
        out[f"{name}_y"] = y.numpy()
        sd = {}
        for grp in ("pre", "post"):
            for i in range(n):
                blk = getattr(model, grp)[i]
                sd[f"{grp}.{i}.weight"] = blk.KL.weight
                sd[f"{grp}.{i}.bias"] = blk.KL.bias
        sd["l.weight"], sd["l.bias"] = model.l.KL.weight, model.l.KL.bias
        for grp in ("c0", "c1"):
            for i in range(n):
                blk = getattr(model, grp)[i]
                sd[f"{grp}.{i}.weight"] = blk.KL.weight.mean(0, keepdim=True)
                sd[f"{grp}.{i}.bias"] = blk.KL.bias.mean(0, keepdim=True)
        for k, v in sd.items():
            out[f"{name}_p_{k}"] = v.detach().numpy()
        # the c0/c1 scalars the reference computed, level by level
        with torch.no_grad():
            im = img[0]
            for i in range(n):
                out[f"{name}_c0_{i}"] = np.array(float(model.c0[i](im)))
                out[f"{name}_c1_{i}"] = np.array(float(model.c1[i](im)))
                im = torch.nn.functional.avg_pool3d(im, (2, 2, 2))
        out[f"{name}_n"] = np.array(n)
    path = os.path.join(HERE, "ref_models.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()
