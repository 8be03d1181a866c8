"""On-disk formats of the reference scenes (mlpcg_amd/formats.py)."""
import struct

import numpy as np
import scipy.sparse as sp

from mlpcg_amd import formats as FM
from mlpcg_amd import scenes as S


def test_flags_roundtrip_and_layout(tmp_path):
    flags = np.array([0, 1, 2, 3, 2, 2], dtype=np.int32)
    p = tmp_path / "flags_1.bin"
    FM.write_flags(p, flags)
    raw = p.read_bytes()
    assert raw[:8] == struct.pack("<Q", 6)
    assert raw[8:] == struct.pack("<6i", *flags)
    assert np.array_equal(FM.read_flags(p), flags)


def test_vector_roundtrip_and_layout(tmp_path):
    v = np.array([0.5, -1.25, 3.0])
    p = tmp_path / "div_v_star_1.bin"
    FM.write_vector(p, v)
    raw = p.read_bytes()
    assert raw == struct.pack("<Q", 3) + struct.pack("<3d", *v)
    assert np.array_equal(FM.load_vector(p), v)
    # the reference's reader reinterprets the size_t header as one double and drops it
    assert np.array_equal(np.fromfile(p, dtype="<f8")[1:], v)


def _struct_writer(A, path, dtype="d"):
    """Byte-for-byte the loop of lib/write_data.py:19-48 (outS = len(indptr))."""
    A = sp.csr_matrix(A)
    with open(path, "wb") as f:
        for t in (A.shape[0], A.shape[1], A.nnz, len(A.indptr), len(A.indptr)):
            f.write(struct.pack("i", t))
        for x in A.data:
            f.write(struct.pack(dtype, x))
        for x in A.indptr:
            f.write(struct.pack("i", x))
        for x in A.indices:
            f.write(struct.pack("i", x))


def test_matrix_both_writer_layouts(tmp_path):
    sc = S.scene_spheres(12, 1, num_spheres=2, rmin=2, rmax=3)
    A = S.build_matrix(sc.flags)
    p1, p2 = tmp_path / "A_eigen.bin", tmp_path / "A_py.bin"
    FM.write_matrix(p1, A, eigen=True)
    _struct_writer(A, p2)
    for p in (p1, p2):
        B = FM.read_matrix(p)
        assert B.shape == A.shape and abs(B - A).max() == 0
    assert p1.stat().st_size == 20 + 8 * A.nnz + 4 * A.shape[0] + 4 * A.nnz
    FM.write_matrix(p1, A, dtype="f")
    assert abs(FM.read_matrix(p1, dtype="f") - A).max() == 0


def test_compression_matches_build_matrix():
    """A full-grid matrix with empty rows/cols for non-fluid cells compresses to the
    fluid-only matrix; vectors compress/expand with the flag mask."""
    sc = S.scene_spheres(10, 2, num_spheres=2, rmin=2, rmax=3)
    A = S.build_matrix(sc.flags)
    idx = sc.fluid_cells
    V = sc.flags.size
    P = sp.csr_matrix((np.ones(idx.size), (idx, np.arange(idx.size))), shape=(V, idx.size))
    full = (P @ A @ P.T).tocsr()
    assert abs(FM.compressed_matrix(full) - A).max() == 0
    v = np.random.default_rng(0).standard_normal(V)
    c = FM.compressed_vector(v, sc.flags)
    assert np.array_equal(c, v[idx])
    e = FM.expand_vector(c, sc.flags)
    assert np.array_equal(e[idx], c) and np.all(np.delete(e, idx) == 0)
