"""The C-ABI library loads (no GPU needed) and exports every symbol include/mlpcg.h
declares, with the argument counts the ctypes binding uses."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mlpcg.h")
LIB = os.path.join(ROOT, "mlpcg_amd", "lib", "libmlpcg_hip.so")


def _declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    decls = {}
    for m in re.finditer(r"\b(mlpcg_\w+)\s*\(([^;{]*?)\)\s*;", txt):
        name, args = m.group(1), m.group(2).strip()
        nargs = 0 if args in ("", "void") else args.count(",") + 1
        decls[name] = nargs
    return decls


def test_header_parses():
    d = _declared()
    assert "mlpcg_sm_block_inference" in d and d["mlpcg_sm_block_inference"] == 13
    assert len(d) >= 30


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_symbol():
    import torch  # noqa: F401  (binds the HIP runtime first, as the package does)
    lib = ctypes.CDLL(LIB)
    for name in _declared():
        assert hasattr(lib, name), name


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_binding_table_matches_header():
    from mlpcg_amd import _lib
    decl = _declared()
    assert set(decl) == set(_lib.SIGNATURES)
    for name, (_, args) in _lib.SIGNATURES.items():
        assert len(args) == decl[name], name


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_error_path_without_gpu():
    """Argument validation fails before any device call, with a message."""
    from mlpcg_amd import _lib
    rc = _lib.lib.mlpcg_stencil_weights(None, 1, None, None, None, 1, 0, 7, 8, 8, 8, None)
    assert rc == -1
    assert "C must be" in _lib.last_error()
