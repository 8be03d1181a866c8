"""The allocation-free SM-linear entry point (partials in the library's device array)
against the workspace entry point, through the C ABI: two calls in a row, the second on a
smaller grid, bit-identical outputs (same kernels, same summation order)."""
import ctypes as C

import pytest
import torch

pytestmark = pytest.mark.gpu


def _call_pair(lib, img, w, b, dt_code, dev):
    from mlpcg_amd._lib import check, ptr, stream_ptr
    Cimg, n1, n2, n3 = (int(s) for s in img.shape)
    acc = torch.float64 if dt_code == 0 else torch.float32
    z_ref = torch.full((1,), float("nan"), dtype=acc, device=dev)
    yb_ref = torch.full((27 * Cimg,), float("nan"), dtype=acc, device=dev)
    z_na = torch.full((1,), float("nan"), dtype=acc, device=dev)
    yb_na = torch.full((27 * Cimg,), float("nan"), dtype=acc, device=dev)
    ws = torch.empty(max(16, int(lib.mlpcg_sm_linear_workspace_size(Cimg, n1, n2, n3, dt_code))),
                     dtype=torch.uint8, device=dev)
    st = stream_ptr(dev)
    check(lib.mlpcg_sm_linear_inference(ptr(img), ptr(w), ptr(b), ptr(z_ref), ptr(yb_ref), ptr(ws), Cimg,
                                        n1, n2, n3, dt_code, st), "sm_linear_inference")
    check(lib.mlpcg_sm_linear_inference_noalloc(ptr(img), ptr(w), ptr(b), ptr(z_na), ptr(yb_na), Cimg, n1, n2,
                                                n3, dt_code, st), "sm_linear_inference_noalloc")
    torch.cuda.synchronize()
    return z_ref, yb_ref, z_na, yb_na


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32, torch.bfloat16])
@pytest.mark.parametrize("Cimg", [1, 3])
def test_noalloc_sm_linear_matches_workspace_entry(dev, dtype, Cimg):
    from mlpcg_amd._lib import lib, code_of
    print(f"noalloc sm_linear dtype={dtype} C={Cimg}", flush=True)
    g = torch.Generator().manual_seed(7 + Cimg)
    acc = torch.float64 if dtype == torch.float64 else torch.float32
    w = (torch.rand(1, Cimg, 3, 3, 3, generator=g, dtype=torch.float64) - 0.5).to(dev, acc)
    b = (torch.rand(1, generator=g, dtype=torch.float64) - 0.5).to(dev, acc)
    # first call: a grid whose reduction uses the full 1024 workgroups; second: smaller
    for shape in ((128, 64, 128), (12, 8, 20)):
        img = torch.rand(Cimg, *shape, generator=g, dtype=torch.float64).to(dev, dtype).contiguous()
        z_ref, yb_ref, z_na, yb_na = _call_pair(lib, img, w, b, code_of(dtype), dev)
        assert torch.isfinite(z_ref).all() and torch.isfinite(yb_ref).all()
        assert torch.equal(z_na, z_ref), (shape, z_na, z_ref)
        assert torch.equal(yb_na, yb_ref), shape


def test_noalloc_sm_linear_refuses_bad_sizes(dev):
    from mlpcg_amd._lib import lib, ptr, stream_ptr, code_of
    img = torch.zeros(1, 8, 8, 8, device=dev)
    w = torch.zeros(1, 1, 3, 3, 3, device=dev)
    b = torch.zeros(1, device=dev)
    z = torch.zeros(1, device=dev)
    yb = torch.zeros(27, device=dev)
    assert int(lib.mlpcg_sm_linear_noalloc_scratch_bytes()) == 1024 * 27 * 3 * 8
    rc = lib.mlpcg_sm_linear_inference_noalloc(ptr(img), ptr(w), ptr(b), ptr(z), ptr(yb), 4, 8, 8, 8,
                                               code_of(torch.float32), stream_ptr(dev))
    assert rc != 0
    rc = lib.mlpcg_sm_linear_inference_noalloc(ptr(img), ptr(w), ptr(b), ptr(z), ptr(yb), 1, 0, 8, 8,
                                               code_of(torch.float32), stream_ptr(dev))
    assert rc != 0
