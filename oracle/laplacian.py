"""Loop-form restatement of the fluid Laplacian (test infrastructure only).

lib/discrete_laplacian.py:122-160 (lap_with_bc) starts from the full 7-point Laplacian
(diagonal 6, off-diagonal -1), lowers the diagonal by one for each solid neighbour
(:148-152, with the box boundary handled the same way through bd, :139-143), then
zeroes the rows and columns of solid (:153-155) and air (:156-159) cells;
compressedMat (lib/read_data.py:164-165) finally drops the empty rows/columns.
This module rebuilds that matrix entry by entry from the flags with a dict-of-keys,
in that order, so it checks mlpcg_amd.scenes.build_matrix by an independent route.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

SOLID_CODES = (0, 1)
FLUID = 2
AIR = 3


def lap_fluid_loops(flags: np.ndarray) -> sp.csr_matrix:
    n1, n2, n3 = flags.shape
    V = n1 * n2 * n3

    def flat(i, j, k):
        return (i * n2 + j) * n3 + k

    entries = {}
    for i in range(n1):
        for j in range(n2):
            for k in range(n3):
                v = flat(i, j, k)
                entries[(v, v)] = 6.0
                for d in ((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)):
                    a, b, c = i + d[0], j + d[1], k + d[2]
                    if 0 <= a < n1 and 0 <= b < n2 and 0 <= c < n3:
                        entries[(v, flat(a, b, c))] = -1.0
                        if flags[a, b, c] in SOLID_CODES:
                            entries[(v, v)] -= 1.0
                    else:
                        # outside the box: a wall, i.e. Neumann like solid
                        entries[(v, v)] -= 1.0
    keep = [flat(i, j, k) for i in range(n1) for j in range(n2) for k in range(n3)
            if flags[i, j, k] == FLUID]
    pos = {v: t for t, v in enumerate(keep)}
    rows, cols, vals = [], [], []
    for (r, c), val in entries.items():
        if r in pos and c in pos and val != 0.0:
            rows.append(pos[r])
            cols.append(pos[c])
            vals.append(val)
    A = sp.coo_matrix((vals, (rows, cols)), shape=(len(keep), len(keep))).tocsr()
    A.sum_duplicates()
    A.sort_indices()
    del V
    return A
