"""Pins of the FP64 solver restatement (oracle/solvers.py)."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from mlpcg_amd import scenes as S
from oracle import solvers as O


def _problem(n=16, seed=0):
    sc = S.scene_spheres(n, seed, num_spheres=2, rmin=2, rmax=4)
    A = S.build_matrix(sc.flags)
    b = S.make_rhs(sc.flags, seed)[0]
    return sc, A, b


def test_npcg_identity_matches_scipy_cg():
    """With P = I, npcg is textbook CG: its iterates equal scipy's (same recurrences up
    to the true-residual recompute, which only changes rounding)."""
    _, A, b = _problem()
    hist_sp = []
    spla.cg(A, b, x0=np.zeros_like(b), rtol=1e-10, atol=0.0, maxiter=40,
            callback=lambda xk: hist_sp.append(np.linalg.norm(b - A @ xk) / np.linalg.norm(b)))
    _, k, hist = O.npcg(b, A, np.zeros_like(b), lambda r: r.copy(), 40, tol=1e-10)
    m = min(len(hist_sp), len(hist) - 1, 30)
    assert m >= 10
    rel = np.abs(np.array(hist[1:m + 1]) - np.array(hist_sp[:m])) / np.array(hist_sp[:m])
    assert rel.max() < 1e-6


def test_npcg_exact_inverse_one_step():
    _, A, b = _problem()
    solve = spla.factorized(A.tocsc())
    x, k, hist = O.npcg(b, A, np.zeros_like(b), solve, 10, tol=1e-10)
    assert k == 1 and hist[-1] < 1e-10


def test_jacobi_pcg_converges_and_flex_equals_fr_for_fixed_spd_p():
    """P = D^-1 (SPD, fixed): Fletcher-Reeves and Polak-Ribiere betas agree in exact
    arithmetic since r_k . z_{k-1} = 0."""
    _, A, b = _problem()
    d = A.diagonal()
    jac = lambda r: r / d
    _, k1, h1 = O.npcg(b, A, np.zeros_like(b), jac, 200, tol=1e-8)
    _, k2, h2 = O.npcg_flex(b, A, np.zeros_like(b), jac, 200, tol=1e-8)
    assert h1[-1] <= 1e-8 and h2[-1] <= 1e-8
    assert abs(k1 - k2) <= 2
    m = min(len(h1), len(h2), 15)
    assert np.allclose(h1[:m], h2[:m], rtol=1e-5)


def test_npsd_monotone_and_converges():
    """npsd minimises the A-norm error along each new direction: the A-norm of the
    error never increases."""
    _, A, b = _problem()
    d = A.diagonal()
    xs = spla.spsolve(A.tocsc(), b)
    errs = []
    x, k, hist = O.npsd(b, A, np.zeros_like(b), lambda r: r / d, 300, tol=1e-8)
    assert hist[-1] < 1e-8
    # rerun recording the A-norm of the error per iteration
    for it in (1, 2, 5, 10, 20):
        xi, _, _ = O.npsd(b, A, np.zeros_like(b), lambda r: r / d, it, tol=0.0)
        e = xi - xs
        errs.append(float(e @ (A @ e)))
    assert all(errs[i + 1] <= errs[i] * (1 + 1e-9) for i in range(len(errs) - 1))


def test_stopping_rule_counts():
    """k counts preconditioner applications; the loop stops at max_it."""
    _, A, b = _problem()
    _, k, hist = O.npcg(b, A, np.zeros_like(b), lambda r: r.copy(), 5, tol=1e-30)
    assert k == 5 and len(hist) == 6
    _, k, hist = O.npsd(b, A, np.zeros_like(b), lambda r: r.copy(), 5, tol=1e-30)
    assert k == 5 and len(hist) == 6


def _nonsymmetric_p(A):
    """A fixed, non-symmetric linear preconditioner: z = D^-1 (r + 0.3 L r) with L the
    strictly lower part of A (the learned SM preconditioner is non-symmetric too)."""
    import scipy.sparse as sp
    d = A.diagonal()
    L = sp.tril(A, k=-1).tocsr()
    M = sp.diags(1.0 / d) @ (sp.identity(A.shape[0]) + 0.3 * L)
    assert abs(M - M.T).max() > 1e-3
    return lambda r: M @ r


def _a_cos(A, u, v):
    return abs(u @ (A @ v)) / np.sqrt((u @ (A @ u)) * (v @ (A @ v)))


def test_npcg_flex_directions_are_a_conjugate_with_nonsymmetric_p():
    """cg_tests.py:159: beta = z_k.(r_k - r_{k-1}) / rz_{k-1}.  With r_k = r_{k-1} -
    alpha A p_{k-1} and alpha = rz_{k-1} / p_{k-1}.A p_{k-1}, this beta is exactly the one
    that makes p_k A-conjugate to p_{k-1} for ANY preconditioner.  Fletcher-Reeves
    (npcg, beta = rz_k / rz_{k-1}) is not, once P is non-symmetric: the check tells the
    two apart."""
    _, A, b = _problem()
    P = _nonsymmetric_p(A)
    dirs = []
    O.npcg_flex(b, A, np.zeros_like(b), P, 25, tol=1e-30, trace=dirs)
    assert len(dirs) == 25
    flex = max(_a_cos(A, dirs[k], dirs[k - 1]) for k in range(1, len(dirs)))
    assert flex < 1e-9, flex
    fr_dirs = []
    O.npcg(b, A, np.zeros_like(b), P, 25, tol=1e-30, trace=fr_dirs)
    fr = max(_a_cos(A, fr_dirs[k], fr_dirs[k - 1]) for k in range(1, len(fr_dirs)))
    assert fr > 1e-3, fr


def test_npsd_directions_a_orthogonal_and_exact_line_search():
    """cg_tests.py:217-281: each new direction Q is A-orthogonal to the (up to two) kept
    directions it was projected against, and the step along Q is an exact line search,
    so the new residual is orthogonal to Q.  In exact arithmetic both hold for any P:
    the kept pair is itself A-orthogonal (by induction), so projecting against the newer
    one first does not spoil the older one.  A sign error in a projection or in the step,
    a projection against the wrong slot, or a dropped projection breaks one of them."""
    _, A, b = _problem()
    P = _nonsymmetric_p(A)
    tr = []
    O.npsd(b, A, np.zeros_like(b), P, 20, tol=1e-30, trace=tr)
    assert len(tr) == 20
    r_prev = b.copy()
    for i, (Q, kept, r_new) in enumerate(tr):
        assert len(kept) == min(i, 2)
        for v in kept:
            assert _a_cos(A, Q, v) < 1e-9, (i, _a_cos(A, Q, v))
        assert abs(Q @ r_new) < 1e-10 * np.linalg.norm(Q) * np.linalg.norm(r_prev), i
        r_prev = r_new
    # without the projections (num_prev = 0: plain steepest descent) consecutive
    # directions are far from A-orthogonal: the orthogonality check has teeth
    tr0 = []
    O.npsd(b, A, np.zeros_like(b), P, 20, tol=1e-30, num_prev=0, trace=tr0)
    worst = max(_a_cos(A, tr0[i][0], tr0[i - 1][0]) for i in range(1, len(tr0)))
    assert worst > 1e-2, worst
