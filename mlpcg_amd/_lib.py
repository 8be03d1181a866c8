"""ctypes binding of libmlpcg_hip.so (the C ABI declared in include/mlpcg.h).

torch supplies device memory and the current HIP stream; every call passes raw
pointers.  torch is imported first so that its HIP runtime (SONAME libamdhip64.so.7)
is the one the library binds to.  There is no fallback: if the library is missing the
import of this module raises.
"""
from __future__ import annotations

import ctypes as C
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("MLPCG_LIB", os.path.join(_HERE, "lib", "libmlpcg_hip.so"))

F64, F32, F16, BF16 = 0, 1, 2, 3
NPCG, NPCG_FLEX, NPSD = 0, 1, 2
MODEL_SMALL_SM, MODEL_SPD, MODEL_SMALL_SPD = 0, 1, 2

DTYPE_CODE = {torch.float64: F64, torch.float32: F32, torch.float16: F16, torch.bfloat16: BF16}
ACC_DTYPE = {F64: torch.float64, F32: torch.float32, F16: torch.float32, BF16: torch.float32}


class MlpcgError(RuntimeError):
    pass


class SolveInfo(C.Structure):
    _fields_ = [("iters", C.c_int), ("status", C.c_int), ("rel_res", C.c_double),
                ("norm_b", C.c_double)]


vp = C.c_void_p
i32 = C.c_int
i64 = C.c_long
f64 = C.c_double
sz = C.c_size_t

# name -> (restype, argtypes); this table is also the list of exported symbols that
# include/mlpcg.h declares (checked by tests/test_capi_symbols.py).
SIGNATURES = {
    "mlpcg_last_error": (C.c_char_p, []),
    "mlpcg_version": (i32, []),
    "mlpcg_sm_block_workspace_size": (sz, [i32, i32, i32, i32, i32, i32]),
    "mlpcg_sm_block_inference": (i32, [vp, vp, vp, vp, vp, vp, i32, i32, i32, i32, i32, i32, vp]),
    "mlpcg_sm_block_forward": (i32, [vp, vp, vp, vp, vp, vp, i32, i32, i32, i32, i32, i32, vp]),
    "mlpcg_sm_block_backward": (i32, [vp, vp, vp, vp, vp, vp, vp, vp, vp, i32, i32, i32, i32, i32,
                                      i32, vp]),
    "mlpcg_sm_block_forward_ex": (i32, [vp, vp, vp, vp, vp, vp, vp, vp, i32, i32, i32, i32, i32, i32, vp]),
    "mlpcg_sm_block_backward_ex": (i32, [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, i32, i32, i32, i32,
                                         i32, i32, vp]),
    "mlpcg_fused_pack_bytes": (sz, [i32]),
    "mlpcg_fused_pack": (i32, [vp, vp, i32, vp, vp]),
    "mlpcg_sm_block_trans_inference": (i32, [vp, vp, vp, vp, vp, vp, i32, i32, i32, i32, i32, i32,
                                             vp]),
    "mlpcg_sm_block_trans_forward": (i32, [vp, vp, vp, vp, vp, vp, i32, i32, i32, i32, i32, i32, vp]),
    "mlpcg_sm_block_trans_backward": (i32, [vp, vp, vp, vp, vp, vp, vp, vp, vp, i32, i32, i32, i32,
                                            i32, i32, vp]),
    "mlpcg_sm_linear_workspace_size": (sz, [i32, i32, i32, i32, i32]),
    "mlpcg_sm_linear_inference": (i32, [vp, vp, vp, vp, vp, vp, i32, i32, i32, i32, i32, vp]),
    "mlpcg_sm_linear_forward": (i32, [vp, vp, vp, vp, vp, vp, i32, i32, i32, i32, i32, vp]),
    "mlpcg_sm_linear_backward": (i32, [vp, vp, vp, vp, i32, i32, vp]),
    "mlpcg_sm_linear_noalloc_scratch_bytes": (sz, []),
    "mlpcg_sm_linear_inference_noalloc": (i32, [vp, vp, vp, vp, vp, i32, i32, i32, i32, i32, vp]),
    "mlpcg_stencil_weights": (i32, [vp, i32, vp, vp, vp, i32, i32, i32, i32, i32, i32, vp]),
    "mlpcg_sm_apply": (i32, [vp, i32, vp, i32, vp, i32, vp, vp, vp, vp, vp, vp, i32, i32, i32, i32,
                             vp]),
    "mlpcg_avgpool2": (i32, [vp, vp, i32, i32, i32, i32, i32, vp]),
    "mlpcg_upsample2": (i32, [vp, vp, i32, i32, i32, i32, i32, f64, vp, vp, vp, i32, vp]),
    "mlpcg_upsample2_adjoint": (i32, [vp, vp, i32, i32, i32, i32, i32, f64, i32, vp]),
    "mlpcg_flags_code": (i32, [vp, vp, i32, i32, i32, vp]),
    "mlpcg_spmv_flags": (i32, [vp, vp, vp, i32, i32, i32, vp]),
    "mlpcg_comb": (i32, [vp, vp, vp, vp, i32, vp, i64, i32, vp]),
    "mlpcg_comb_backward": (i32, [vp, vp, vp, vp, vp, i32, vp, vp, i64, i32, vp, i32, vp, vp]),
    "mlpcg_loss_resid": (i32, [vp, i32, vp, vp, vp, i32, i32, i32, vp, vp, vp]),
    "mlpcg_loss_grad": (i32, [vp, vp, vp, vp, i32, i32, i32, i32, vp]),
    "mlpcg_spmv_csr": (i32, [vp, vp, vp, vp, vp, i64, i32, vp]),
    "mlpcg_dot_workspace_size": (sz, []),
    "mlpcg_dot": (i32, [vp, vp, i64, vp, vp, vp]),
    "mlpcg_dot_f64_workspace_size": (sz, []),
    "mlpcg_dot_f64": (i32, [vp, vp, i64, vp, vp, vp]),
    "mlpcg_lanczos_update": (i32, [vp, vp, vp, vp, vp, i64, vp]),
    "mlpcg_normalize_f64": (i32, [vp, vp, vp, vp, i64, vp]),
    "mlpcg_precond_create": (i32, [C.POINTER(vp), i32, i32, i32, i32, i32, i32, i32, i32, i32, i32,
                                   vp, C.POINTER(vp), C.POINTER(vp), vp]),
    "mlpcg_precond_update": (i32, [vp, vp, C.POINTER(vp), C.POINTER(vp), vp]),
    "mlpcg_precond_apply": (i32, [vp, vp, vp, vp, vp, vp]),
    "mlpcg_precond_bytes": (sz, [vp]),
    "mlpcg_precond_destroy": (None, [vp]),
    "mlpcg_precond_create_ex": (i32, [C.POINTER(vp), i32, i32, i32, i32, i32, i32, i32, i32, i32, i32,
                                      i32, i32, vp, C.POINTER(vp), C.POINTER(vp), vp]),
    "mlpcg_precond_fused_blocks": (i32, [vp]),
    "mlpcg_precond_coarse_tail": (i32, [vp]),
    "mlpcg_precond_apply_block": (i32, [vp, i32, vp, i32, vp, i32, vp, vp]),
    "mlpcg_solver_create_flags": (i32, [C.POINTER(vp), vp, i32, i32, i32, vp]),
    "mlpcg_solver_create_csr": (i32, [C.POINTER(vp), vp, vp, vp, i64, vp, i32, i32, i32, vp]),
    "mlpcg_solver_solve": (i32, [vp, vp, vp, vp, vp, i32, i32, f64, f64, vp, C.POINTER(SolveInfo),
                                 vp]),
    "mlpcg_solver_set_graphs": (i32, [vp, i32]),
    "mlpcg_solver_destroy": (None, [vp]),
}


def _load():
    if not os.path.exists(LIB_PATH):
        raise MlpcgError(f"HIP library not built: {LIB_PATH} (run python -m mlpcg_amd.build)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def last_error() -> str:
    msg = lib.mlpcg_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str = ""):
    if rc != 0:
        raise MlpcgError(f"{what}: {last_error()}")


def ptr(t) -> int | None:
    """Device pointer of a tensor (None for None)."""
    if t is None:
        return None
    return t.data_ptr()


def stream_ptr(device=None) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def code_of(dtype: torch.dtype) -> int:
    try:
        return DTYPE_CODE[dtype]
    except KeyError as e:
        raise MlpcgError(f"unsupported dtype {dtype}") from e


def require_cuda(*tensors):
    for t in tensors:
        if t is None:
            continue
        if not t.is_cuda:
            raise MlpcgError("tensor must be on the GPU (HIP device)")
        if not t.is_contiguous():
            raise MlpcgError("tensor must be contiguous")
