"""Synthetic scenes: cell flags, the fluid Laplacian, the image channels and the RHS.

Flag codes follow lib/read_data.py:13-16 (BOUNDARY 0, SOLID 1, FLUID 2, AIR 3).
The matrix follows lib/discrete_laplacian.py:122-160 restricted to fluid rows and
columns as lib/read_data.py:164-165 (compressedMat) leaves it:

* A_ii = number of face neighbours that are not solid (fluid or air); a neighbour
  outside the grid counts as solid (Neumann);
* A_ij = -1 for every fluid face neighbour j;
* rows and columns of solid and air cells are dropped, and the unknowns are the
  fluid cells in flattened (z-fastest) order, as compressedVec (:166-168) selects them.

Grids are (N1, N2, N3) with the last axis fastest; "z" (the vertical axis used for
water levels and pillars) is the last axis.

Everything here is host-side setup (numpy / scipy); it never runs per iteration.
"""
from __future__ import annotations

import dataclasses
import numpy as np
import scipy.sparse as sp
import scipy.ndimage as ndi

BOUNDARY = 0
SOLID = 1
FLUID = 2
AIR = 3

# Face-neighbour offsets in (axis, step) form, shared by the matrix builder and the
# matrix-free operator.
FACE_OFFSETS = ((0, -1), (0, 1), (1, -1), (1, 1), (2, -1), (2, 1))


@dataclasses.dataclass
class Scene:
    """One Poisson problem on a regular grid."""
    flags: np.ndarray          # (N1, N2, N3) int8 flag codes
    seed: int
    name: str

    @property
    def shape(self):
        return tuple(self.flags.shape)

    @property
    def fluid_cells(self) -> np.ndarray:
        """Flat indices of fluid cells (np.argwhere(flags == FLUID), test.py:61)."""
        return np.flatnonzero(self.flags.ravel() == FLUID)


def _walls(n1, n2, n3):
    flags = np.full((n1, n2, n3), FLUID, dtype=np.int8)
    flags[0, :, :] = flags[-1, :, :] = BOUNDARY
    flags[:, 0, :] = flags[:, -1, :] = BOUNDARY
    flags[:, :, 0] = flags[:, :, -1] = BOUNDARY
    return flags


def _paint_sphere(flags, center, radius, value=SOLID):
    c = np.asarray(center, dtype=np.float64)
    lo = np.maximum(np.floor(c - radius).astype(int), 0)
    hi = np.minimum(np.ceil(c + radius).astype(int) + 1, flags.shape)
    if np.any(hi <= lo):
        return
    g = np.ogrid[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
    d2 = (g[0] - c[0]) ** 2 + (g[1] - c[1]) ** 2 + (g[2] - c[2]) ** 2
    sub = flags[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
    sub[(d2 <= radius * radius) & (sub != BOUNDARY)] = value


def _paint_pillar(flags, cx, cy, radius, value=SOLID):
    """Vertical (along the last axis) solid cylinder through the whole grid height."""
    n1, n2, _ = flags.shape
    lo0, hi0 = max(int(np.floor(cx - radius)), 0), min(int(np.ceil(cx + radius)) + 1, n1)
    lo1, hi1 = max(int(np.floor(cy - radius)), 0), min(int(np.ceil(cy + radius)) + 1, n2)
    g0, g1 = np.ogrid[lo0:hi0, lo1:hi1]
    disk = (g0 - cx) ** 2 + (g1 - cy) ** 2 <= radius * radius
    sub = flags[lo0:hi0, lo1:hi1, :]
    m = disk[:, :, None] & (sub != BOUNDARY)
    sub[m] = value


def _set_air_above(flags, level):
    """Cells with z index > level (and not boundary/solid) become air."""
    n3 = flags.shape[2]
    z0 = int(np.floor(level)) + 1
    if z0 >= n3:
        return
    sub = flags[:, :, z0:]
    sub[sub == FLUID] = AIR


def _fix_isolated(flags):
    """Fluid cells whose six neighbours are all solid make zero rows of A, which the
    reference's compressedMat drops while compressedVec keeps (lib/read_data.py:164-168).
    They are reclassified as solid so rows and unknowns stay aligned."""
    nonsolid = (flags >= FLUID)
    cnt = np.zeros(flags.shape, dtype=np.int16)
    for ax, st in FACE_OFFSETS:
        cnt += _shift(nonsolid, ax, st)
    iso = (flags == FLUID) & (cnt == 0)
    flags[iso] = SOLID
    return flags


def _shift(a, axis, step):
    """out[v] = a[v + step * e_axis], zero (False) outside the grid."""
    out = np.zeros_like(a)
    src = [slice(None)] * 3
    dst = [slice(None)] * 3
    n = a.shape[axis]
    if step > 0:
        src[axis] = slice(step, n)
        dst[axis] = slice(0, n - step)
    else:
        src[axis] = slice(0, n + step)
        dst[axis] = slice(-step, n)
    out[tuple(dst)] = a[tuple(src)]
    return out


def scene_sphere64(seed: int = 0) -> Scene:
    """64^3: walls one cell thick, one solid sphere r=10 at (32,32,24), air for z>48."""
    flags = _walls(64, 64, 64)
    _paint_sphere(flags, (32, 32, 24), 10)
    _set_air_above(flags, 48)
    return Scene(_fix_isolated(flags), seed, "sphere64")


def scene_spheres(n: int, seed: int, num_spheres: int = 6, rmin: float = 6, rmax: float = 20,
                  level_frac=(0.4, 0.7), air_frac=None, name="spheres") -> Scene:
    """Walls + seeded random solid spheres + air above a seeded water level.

    With ``air_frac=(lo, hi)`` the water level is (1 - U(lo, hi)) * n instead of
    U(level_frac) * n.
    """
    rng = np.random.default_rng(seed)
    flags = _walls(n, n, n)
    for _ in range(num_spheres):
        r = rng.uniform(rmin, rmax)
        c = rng.uniform(1, n - 1, size=3)
        _paint_sphere(flags, c, r)
    if air_frac is not None:
        level = (1.0 - rng.uniform(*air_frac)) * n
    else:
        level = rng.uniform(*level_frac) * n
    _set_air_above(flags, level)
    return Scene(_fix_isolated(flags), seed, f"{name}{n}")


def scene_pillars(n: int, seed: int, num_pillars: int = 24, rmin: float = 4, rmax: float = 10,
                  level_frac: float = 0.55) -> Scene:
    """Walls + seeded vertical solid pillars + air above z = level_frac * n."""
    rng = np.random.default_rng(seed)
    flags = _walls(n, n, n)
    for _ in range(num_pillars):
        r = rng.uniform(rmin, rmax)
        cx, cy = rng.uniform(1, n - 1, size=2)
        _paint_pillar(flags, cx, cy, r)
    _set_air_above(flags, level_frac * n)
    return Scene(_fix_isolated(flags), seed, f"pillars{n}")


def scene_for_config(idx: int, seed: int = 0, n: int | None = None) -> Scene:
    """Scene generator per workload: 0 = 64^3 sphere, 1/2 = 128^3 six spheres,
    3 = 256^3 pillars, 4 = 256^3 forty spheres (air fraction 20-50 %)."""
    if idx == 0:
        return scene_sphere64(seed)
    if idx in (1, 2):
        return scene_spheres(n or 128, seed)
    if idx == 3:
        return scene_pillars(n or 256, seed)
    if idx == 4:
        return scene_spheres(n or 256, seed, num_spheres=40, rmin=4, rmax=30,
                             air_frac=(0.2, 0.5))
    raise ValueError(f"unknown workload index {idx}")


def neighbor_counts(flags: np.ndarray):
    """(#non-solid neighbours, #fluid neighbours) per cell; outside the grid = solid."""
    nonsolid = flags >= FLUID
    fluid = flags == FLUID
    ns = np.zeros(flags.shape, dtype=np.int16)
    fl = np.zeros(flags.shape, dtype=np.int16)
    for ax, st in FACE_OFFSETS:
        ns += _shift(nonsolid, ax, st)
        fl += _shift(fluid, ax, st)
    return ns, fl


def build_matrix(flags: np.ndarray, dtype=np.float64) -> sp.csr_matrix:
    """Fluid-only 7-point Laplacian in CSR with sorted column indices."""
    shape = flags.shape
    flat = flags.ravel()
    fluid_idx = np.flatnonzero(flat == FLUID)
    n = fluid_idx.size
    comp = np.full(flat.size, -1, dtype=np.int64)
    comp[fluid_idx] = np.arange(n)
    ns, _ = neighbor_counts(flags)
    rows = [np.arange(n)]
    cols = [np.arange(n)]
    vals = [ns.ravel()[fluid_idx].astype(dtype)]
    coords = np.unravel_index(fluid_idx, shape)
    strides = (shape[1] * shape[2], shape[2], 1)
    for ax, st in FACE_OFFSETS:
        c = coords[ax] + st
        inside = (c >= 0) & (c < shape[ax])
        nb = fluid_idx[inside] + st * strides[ax]
        is_fluid = flat[nb] == FLUID
        r = np.flatnonzero(inside)[is_fluid]
        rows.append(r)
        cols.append(comp[nb[is_fluid]])
        vals.append(np.full(r.size, -1.0, dtype=dtype))
    A = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(n, n)).tocsr()
    A.sort_indices()
    return A


def image_channels(flags: np.ndarray, num_imgs: int, dtype=np.float32) -> np.ndarray:
    """Image input of the preconditioner, shape (C, N1, N2, N3).

    C = 1: diag(A)/6 on fluid cells, 0 elsewhere.
    C = 3: binary [air, fluid, non-solid] as lib/read_data.py:35-39.
    C = 2: binary [air, non-solid] as lib/read_data.py:31-34.
    """
    if num_imgs == 1:
        ns, _ = neighbor_counts(flags)
        img = np.where(flags == FLUID, ns / 6.0, 0.0)
        return img[None].astype(dtype)
    air = (flags == AIR)
    fluid = (flags == FLUID)
    nonsolid = (flags > SOLID)
    if num_imgs == 2:
        return np.stack([air, nonsolid]).astype(dtype)
    if num_imgs == 3:
        return np.stack([air, fluid, nonsolid]).astype(dtype)
    raise ValueError("num_imgs must be 1, 2 or 3")


def fluid_components_without_air(flags: np.ndarray):
    """Labels of 6-connected fluid components that touch no air cell."""
    fluid = flags == FLUID
    lab, nlab = ndi.label(fluid)
    air = flags == AIR
    touch = np.zeros(flags.shape, dtype=bool)
    for ax, st in FACE_OFFSETS:
        touch |= _shift(air, ax, st)
    touched = np.unique(lab[fluid & touch])
    closed = np.setdiff1d(np.arange(1, nlab + 1), touched)
    return lab, closed


def make_rhs(flags: np.ndarray, seed: int, count: int = 1, dtype=np.float64) -> np.ndarray:
    """Seeded N(0,1) values on fluid cells, mean-subtracted per fluid component with
    no air contact (those components make A singular).  Returns (count, n_fluid)."""
    rng = np.random.default_rng(seed)
    fluid_idx = np.flatnonzero(flags.ravel() == FLUID)
    b = rng.standard_normal((count, fluid_idx.size))
    lab, closed = fluid_components_without_air(flags)
    if closed.size:
        lab_f = lab.ravel()[fluid_idx]
        for c in closed:
            m = lab_f == c
            b[:, m] -= b[:, m].mean(axis=1, keepdims=True)
    return b.astype(dtype)


def expand(vec: np.ndarray, flags: np.ndarray) -> np.ndarray:
    """Scatter a fluid-cell vector onto the full grid (lib/read_data.py:169-174)."""
    out = np.zeros(flags.size, dtype=vec.dtype)
    out[flags.ravel() == FLUID] = vec
    return out.reshape(flags.shape)
