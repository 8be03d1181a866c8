"""Host logic: scene generators, the fluid Laplacian and the image channels."""
import numpy as np
import pytest
import scipy.sparse as sp

from mlpcg_amd import scenes as S
from oracle.laplacian import lap_fluid_loops


def _small_flags(seed, n=6):
    rng = np.random.default_rng(seed)
    f = rng.choice([S.SOLID, S.FLUID, S.FLUID, S.FLUID, S.AIR], size=(n, n + 1, n - 1)).astype(np.int8)
    f[0] = S.BOUNDARY
    return S._fix_isolated(f)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_matrix_vs_loop_restatement(seed):
    flags = _small_flags(seed)
    A = S.build_matrix(flags)
    B = lap_fluid_loops(flags)
    assert A.shape == B.shape
    assert abs(A - B).max() == 0


def test_matrix_hand_case():
    """3x3x3 box of walls with one fluid cell in the middle next to one air cell."""
    f = np.full((3, 3, 3), S.BOUNDARY, dtype=np.int8)
    f[1, 1, 1] = S.FLUID
    f[1, 1, 2] = S.AIR
    A = S.build_matrix(f).toarray()
    assert A.tolist() == [[1.0]]  # one non-solid neighbour (air), no fluid neighbour
    f[1, 2, 1] = S.FLUID
    A = S.build_matrix(f).toarray()
    # cell (1,1,1): air + fluid neighbours -> 2; cell (1,2,1): fluid neighbour -> 1
    assert A.tolist() == [[2.0, -1.0], [-1.0, 1.0]]


def test_matrix_properties_on_scene():
    sc = S.scene_sphere64()
    A = S.build_matrix(sc.flags)
    assert abs(A - A.T).max() == 0
    ns, fl = S.neighbor_counts(sc.flags)
    fluid = sc.flags.ravel() == S.FLUID
    air = (sc.flags == S.AIR)
    n_air = np.zeros(sc.flags.shape, dtype=np.int16)
    for ax, st in S.FACE_OFFSETS:
        n_air += S._shift(air, ax, st)
    rowsum = np.asarray(A.sum(axis=1)).ravel()
    assert np.array_equal(rowsum, n_air.ravel()[fluid])  # row sum = #air neighbours >= 0
    d = A.diagonal()
    assert d.min() >= 1 and d.max() <= 6
    assert A.shape[0] == int(fluid.sum())


def test_sphere64_geometry():
    sc = S.scene_sphere64()
    f = sc.flags
    assert f.shape == (64, 64, 64)
    assert np.all(f[0] == S.BOUNDARY) and np.all(f[:, :, -1] == S.BOUNDARY)
    assert f[32, 32, 24] == S.SOLID and f[32, 32, 13] == S.FLUID
    assert f[10, 10, 49] == S.AIR and f[10, 10, 48] == S.FLUID


def test_generators_are_seeded():
    a = S.scene_for_config(1, seed=5, n=64).flags
    b = S.scene_for_config(1, seed=5, n=64).flags
    c = S.scene_for_config(1, seed=6, n=64).flags
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    p = S.scene_for_config(3, seed=1, n=64).flags
    assert (p == S.AIR).any() and (p == S.SOLID).any()
    # pillars are vertical: the solid set is the same in every interior z-slice below air
    assert np.array_equal(p[:, :, 5] == S.SOLID, p[:, :, 20] == S.SOLID)


def test_rhs_consistent_on_closed_components():
    f = np.full((8, 8, 8), S.BOUNDARY, dtype=np.int8)
    f[1:7, 1:7, 1:7] = S.FLUID
    f[1:7, 4, 1:7] = S.SOLID           # wall splitting the box in two
    f[1:7, 5:7, 6] = S.AIR             # only the right part touches air
    b = S.make_rhs(f, seed=3)[0]
    lab, closed = S.fluid_components_without_air(f)
    assert closed.size == 1
    lab_f = lab.ravel()[f.ravel() == S.FLUID]
    assert abs(b[lab_f == closed[0]].sum()) < 1e-12
    A = S.build_matrix(f)
    # A is singular on the closed part; b is in its range: the Neumann null vector
    ones = (lab_f == closed[0]).astype(float)
    assert np.allclose(A @ ones, 0.0)
    assert abs(ones @ b) < 1e-12


def test_image_channels():
    sc = S.scene_sphere64()
    img1 = S.image_channels(sc.flags, 1)
    A = S.build_matrix(sc.flags)
    fluid = sc.flags.ravel() == S.FLUID
    assert np.allclose(img1.ravel()[fluid], A.diagonal() / 6.0)
    assert np.all(img1.ravel()[~fluid] == 0)
    img3 = S.image_channels(sc.flags, 3)
    assert img3.shape == (3, 64, 64, 64)
    assert np.array_equal(img3[1] == 1, sc.flags == S.FLUID)
    assert np.array_equal(img3[2] == 1, sc.flags > S.SOLID)
