"""Readers/writers of the reference's on-disk scene data.

* flags_*.bin       size_t N, then N int32 flag codes      (lib/read_data.py:18-26)
* div_v_star_*.bin  size_t N, then N float64 values         (lib/read_data.py:41-48,
                    cxx_src/lib/BinaryIO.h:52-58, lib/write_data.py:11-17)
* A_*.bin           int32 rows, cols, nnz, outS, innS; nnz values (float64 'd' or
                    float32 'f'); outS int32 outer (row) pointers; nnz int32 inner (column)
                    indices (lib/read_data.py:110-162, cxx_src/lib/BinaryIO.h:206-231).
                    Eigen writes outS = rows (no trailing nnz); lib/write_data.py:19-48
                    writes outS = rows + 1; both read back the same.
* compressed_matrix / compressed_vector: restriction to the fluid unknowns
  (lib/read_data.py:164-174).

All I/O is vectorised numpy (no per-element struct calls).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from .scenes import FLUID

_SIZE_T = np.dtype("<u8")


def read_flags(path) -> np.ndarray:
    with open(path, "rb") as f:
        n = int(np.frombuffer(f.read(8), dtype=_SIZE_T)[0])
        data = np.frombuffer(f.read(4 * n), dtype="<i4")
    if data.size != n:
        raise ValueError(f"{path}: expected {n} flags, found {data.size}")
    return data.astype(np.int32)


def write_flags(path, flags) -> None:
    flags = np.ascontiguousarray(np.asarray(flags).ravel(), dtype="<i4")
    with open(path, "wb") as f:
        f.write(np.array([flags.size], dtype=_SIZE_T).tobytes())
        f.write(flags.tobytes())


def load_vector(path, normalize: bool = False) -> np.ndarray:
    with open(path, "rb") as f:
        n = int(np.frombuffer(f.read(8), dtype=_SIZE_T)[0])
        v = np.frombuffer(f.read(8 * n), dtype="<f8").copy()
    if v.size != n:
        raise ValueError(f"{path}: expected {n} values, found {v.size}")
    return v / np.linalg.norm(v) if normalize else v


def write_vector(path, v) -> None:
    v = np.ascontiguousarray(np.asarray(v).ravel(), dtype="<f8")
    with open(path, "wb") as f:
        f.write(np.array([v.size], dtype=_SIZE_T).tobytes())
        f.write(v.tobytes())


def read_matrix(path, dtype: str = "d", shape=None) -> sp.csr_matrix:
    vt = np.dtype("<f8") if dtype == "d" else np.dtype("<f4")
    with open(path, "rb") as f:
        rows, cols, nnz, out_s, _inn_s = (int(t) for t in np.frombuffer(f.read(20), dtype="<i4"))
        vals = np.frombuffer(f.read(vt.itemsize * nnz), dtype=vt)
        outer = np.frombuffer(f.read(4 * out_s), dtype="<i4")
        inner = np.frombuffer(f.read(4 * nnz), dtype="<i4")
    if vals.size != nnz or inner.size != nnz or outer.size != out_s:
        raise ValueError(f"{path}: truncated matrix file")
    indptr = np.empty(rows + 1, dtype=np.int64)
    indptr[:min(rows, out_s)] = outer[:rows]
    indptr[rows] = nnz
    A = sp.csr_matrix((vals.astype(np.float64 if dtype == "d" else np.float32), inner.astype(np.int64),
                       indptr), shape=shape or (rows, cols))
    return A


def write_matrix(path, A, dtype: str = "d", eigen: bool = True) -> None:
    """eigen=True: outS = rows (Eigen's Serialize); False: outS = rows + 1 (write_data.py)."""
    A = sp.csr_matrix(A)
    A.sort_indices()
    vt = np.dtype("<f8") if dtype == "d" else np.dtype("<f4")
    rows, cols = A.shape
    outer = A.indptr[:rows] if eigen else A.indptr
    out_s = outer.size
    with open(path, "wb") as f:
        f.write(np.array([rows, cols, A.nnz, out_s, out_s], dtype="<i4").tobytes())
        f.write(np.ascontiguousarray(A.data, dtype=vt).tobytes())
        f.write(np.ascontiguousarray(outer, dtype="<i4").tobytes())
        f.write(np.ascontiguousarray(A.indices, dtype="<i4").tobytes())


def compressed_matrix(A, flags=None) -> sp.csr_matrix:
    """Drop empty rows and columns (lib/read_data.py:164-165)."""
    A = sp.csr_matrix(A)
    return A[A.getnnz(1) > 0][:, A.getnnz(0) > 0]


def compressed_vector(b, flags) -> np.ndarray:
    """Values at the fluid cells (lib/read_data.py:166-168)."""
    return np.asarray(b)[np.flatnonzero(np.asarray(flags).ravel() == FLUID)]


def expand_vector(b, flags) -> np.ndarray:
    """Inverse of compressed_vector with zeros elsewhere (lib/read_data.py:169-174)."""
    flags = np.asarray(flags).ravel()
    v = np.zeros(flags.shape, dtype=np.asarray(b).dtype)
    v[flags == FLUID] = b
    return v
