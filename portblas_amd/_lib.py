"""ctypes binding of libportblas_amd.so (the C ABI declared in include/portblas_amd.h).

The library is loaded lazily.  If PyTorch is importable it is imported first,
so the process has exactly one HIP runtime (torch ships libamdhip64.so.7 with
the same SONAME the library links against).  A missing library is an error:
there is no fallback implementation.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libportblas_amd.so"

_p = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int
_f32 = ctypes.c_float
_f64 = ctypes.c_double
_ch = ctypes.c_char
_sz = ctypes.c_size_t

# name -> argtypes (restype is pb_status_t = int unless listed in _RESTYPES)
SIGNATURES: dict[str, list] = {
    "pb_handle_create": [_p, ctypes.POINTER(_p)],
    "pb_handle_destroy": [_p],
    "pb_handle_set_stream": [_p, _p],
    "pb_handle_get_stream": [_p],
    "pb_handle_get_num_compute_units": [_p],
    "pb_handle_wait": [_p],
    "pb_handle_reserve_workspace": [_p, _sz],
    "pb_handle_workspace_bytes": [_p],
    "pb_handle_wait_event": [_p, _p],
    "pb_handle_record_event": [_p, _p],
    "pb_status_string": [_i32],
    "pb_last_error": [],
    "pb_abi_version": [],
    "pb_sgemm": [_p, _ch, _ch, _i64, _i64, _i64, _f32, _p, _i64, _p, _i64, _f32, _p, _i64],
    "pb_dgemm": [_p, _ch, _ch, _i64, _i64, _i64, _f64, _p, _i64, _p, _i64, _f64, _p, _i64],
    "pb_sgemm_noalloc": [_p, _ch, _ch, _i64, _i64, _i64, _f32, _p, _i64, _p, _i64, _f32, _p, _i64],
    "pb_sgemm_batched": [_p, _ch, _ch, _i64, _i64, _i64, _f32, _p, _i64, _p, _i64, _f32, _p, _i64, _i64, _i32],
    "pb_dgemm_batched": [_p, _ch, _ch, _i64, _i64, _i64, _f64, _p, _i64, _p, _i64, _f64, _p, _i64, _i64, _i32],
    "pb_sgemm_strided_batched": [_p, _ch, _ch, _i64, _i64, _i64, _f32, _p, _i64, _i64, _p, _i64, _i64,
                                 _f32, _p, _i64, _i64, _i64],
    "pb_dgemm_strided_batched": [_p, _ch, _ch, _i64, _i64, _i64, _f64, _p, _i64, _i64, _p, _i64, _i64,
                                 _f64, _p, _i64, _i64, _i64],
    "pb_gemm_ex": [_p, _i32, _ch, _ch, _i64, _i64, _i64, _f32, _p, _i64, _i64, _p, _i64, _i64, _f32, _p,
                   _i64, _i64, _i64],
    "pb_gemm_bf16": [_p, _ch, _ch, _i64, _i64, _i64, _f32, _p, _i64, _i64, _p, _i64, _i64, _f32, _p,
                     _i64, _i64, _i64],
    "pb_pack_bf16x3": [_p, _i32, _ch, _i64, _i64, _p, _i64, _p, _p, _i64],
    "pb_gemm_bf16x3_packed": [_p, _i64, _i64, _i64, _f32, _p, _p, _i64, _p, _p, _i64, _f32, _p, _i64],
    "pb_pack_fp16x3": [_p, _i32, _ch, _i64, _i64, _p, _i64, _p, _p, _i64, _p],
    "pb_gemm_fp16x3_packed": [_p, _i64, _i64, _i64, _f32, _p, _p, _p, _i64, _p, _p, _p, _i64, _f32, _p, _i64],
    "pb_gemm_grouped_plan_create": [_p, _i32, _p, _i64, ctypes.POINTER(_p)],
    "pb_gemm_grouped_plan_create_ex": [_p, _i32, _p, _i64, _p, ctypes.POINTER(_p)],
    "pb_gemm_grouped_plan_run": [_p, _p],
    "pb_gemm_grouped_plan_destroy": [_p],
    "pb_gemm_grouped": [_p, _i32, _p, _i64],
    "pb_hgemm": [_p, _ch, _ch, _i64, _i64, _i64, _f32, _p, _i64, _p, _i64, _f32, _p, _i64],
    "pb_gemm_f16": [_p, _ch, _ch, _i64, _i64, _i64, _f32, _p, _i64, _i64, _p, _i64, _i64, _f32, _p,
                    _i64, _i64, _i64],
    "pb_cgemm": [_p, _ch, _ch, _i64, _i64, _i64, _p, _p, _i64, _p, _i64, _p, _p, _i64],
    "pb_zgemm": [_p, _ch, _ch, _i64, _i64, _i64, _p, _p, _i64, _p, _i64, _p, _p, _i64],
    "pb_ssymm": [_p, _ch, _ch, _i64, _i64, _f32, _p, _i64, _p, _i64, _f32, _p, _i64],
    "pb_dsymm": [_p, _ch, _ch, _i64, _i64, _f64, _p, _i64, _p, _i64, _f64, _p, _i64],
    "pb_strsm": [_p, _ch, _ch, _ch, _ch, _i64, _i64, _f32, _p, _i64, _p, _i64],
    "pb_dtrsm": [_p, _ch, _ch, _ch, _ch, _i64, _i64, _f64, _p, _i64, _p, _i64],
    "pb_sgemv": [_p, _ch, _i64, _i64, _f32, _p, _i64, _p, _i64, _f32, _p, _i64],
    "pb_dgemv": [_p, _ch, _i64, _i64, _f64, _p, _i64, _p, _i64, _f64, _p, _i64],
    "pb_sdot": [_p, _i64, _p, _i64, _p, _i64, _p],
    "pb_ddot": [_p, _i64, _p, _i64, _p, _i64, _p],
    "pb_sasum": [_p, _i64, _p, _i64, _p],
    "pb_dasum": [_p, _i64, _p, _i64, _p],
    "pb_snrm2": [_p, _i64, _p, _i64, _p],
    "pb_dnrm2": [_p, _i64, _p, _i64, _p],
    "pb_sdsdot": [_p, _i64, _f32, _p, _i64, _p, _i64, _p],
    "pb_sdsdot_sync": [_p, _i64, _f32, _p, _i64, _p, _i64, ctypes.POINTER(_f32)],
    "pb_saxpy": [_p, _i64, _f32, _p, _i64, _p, _i64],
    "pb_daxpy": [_p, _i64, _f64, _p, _i64, _p, _i64],
    "pb_sscal": [_p, _i64, _f32, _p, _i64],
    "pb_dscal": [_p, _i64, _f64, _p, _i64],
    "pb_scopy": [_p, _i64, _p, _i64, _p, _i64],
    "pb_dcopy": [_p, _i64, _p, _i64, _p, _i64],
    "pb_sswap": [_p, _i64, _p, _i64, _p, _i64],
    "pb_dswap": [_p, _i64, _p, _i64, _p, _i64],
    "pb_isamax": [_p, _i64, _p, _i64, _p],
    "pb_isamin": [_p, _i64, _p, _i64, _p],
    "pb_idamax": [_p, _i64, _p, _i64, _p],
    "pb_idamin": [_p, _i64, _p, _i64, _p],
    "pb_isamax_64": [_p, _i64, _p, _i64, _p],
    "pb_isamin_64": [_p, _i64, _p, _i64, _p],
    "pb_sdot_sync": [_p, _i64, _p, _i64, _p, _i64, ctypes.POINTER(_f32)],
    "pb_sasum_sync": [_p, _i64, _p, _i64, ctypes.POINTER(_f32)],
    "pb_snrm2_sync": [_p, _i64, _p, _i64, ctypes.POINTER(_f32)],
    "pb_isamax_sync": [_p, _i64, _p, _i64, ctypes.POINTER(ctypes.c_int32)],
    "pb_isamin_sync": [_p, _i64, _p, _i64, ctypes.POINTER(ctypes.c_int32)],
    "pb_sreduction": [_p, _i32, _i32, _p, _i64, _p, _i64, _i64],
    "pb_dreduction": [_p, _i32, _i32, _p, _i64, _p, _i64, _i64],
    "pb_graph_capture_begin": [_p],
    "pb_graph_capture_end": [_p, ctypes.POINTER(_p)],
    "pb_graph_launch": [_p, _p],
    "pb_graph_destroy": [_p],
    "pb_tune_num_configs": [],
    "pb_tune_config_name": [_i32],
    "pb_tune_time_config": [_p, _i32, _ch, _ch, _i64, _i64, _i64, _p, _i64, _p, _i64, _p, _i64, _i64,
                            _i32, _i32, ctypes.POINTER(_f32)],
    "pb_gemm_bf16_config": [_p, _i32, _ch, _ch, _i64, _i64, _i64, _f32, _p, _i64, _i64, _p, _i64, _i64,
                            _f32, _p, _i64, _i64, _i64],
    "pb_tune_set_selector": [_p, _i32, _p, _p, _p, _p, _p],
    "pb_tune_select": [_p, _ch, _ch, _i64, _i64, _i64, _i64],
    "pb_tune_num_plain_configs": [],
    "pb_tune_plain_config_name": [_i32],
    "pb_gemm_plain_config": [_p, _i32, _i32, _ch, _ch, _i64, _i64, _i64, _f64, _p, _i64, _p, _i64, _f64, _p,
                             _i64, _i64],
    "pb_tune_time_plain": [_p, _i32, _i32, _ch, _ch, _i64, _i64, _i64, _p, _i64, _p, _i64, _p, _i64, _i64,
                           _i32, _i32, ctypes.POINTER(_f32)],
    "pb_tune_set_plain_selector": [_p, _i32, _p, _p, _p, _p, _p],
    "pb_tune_select_plain": [_p, _i32, _ch, _ch, _i64, _i64, _i64, _i64, _i64],
    "pb_mlp_forward": [_p, _i32, _p, _p, _p, _p, _i64],
    "pb_mlp_forward_bf16": [_p, _i32, _p, _p, _p, _p, _p, _i64],
    "pb_mlp_backward": [_p, _i32, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i64],
    "pb_mlp_adam": [_p, _p, _p, _p, _p, _i64, _f32, _f32, _f32, _f32, _i32],
}

_RESTYPES = {
    "pb_handle_get_stream": _p,
    "pb_handle_get_num_compute_units": _i32,
    "pb_handle_workspace_bytes": _sz,
    "pb_status_string": ctypes.c_char_p,
    "pb_last_error": ctypes.c_char_p,
    "pb_abi_version": _i32,
    "pb_tune_num_configs": _i32,
    "pb_tune_config_name": ctypes.c_char_p,
    "pb_tune_select": _i32,
    "pb_tune_num_plain_configs": _i32,
    "pb_tune_plain_config_name": ctypes.c_char_p,
    "pb_tune_select_plain": _i32,
}

_lib = None


def declared_symbols() -> list[str]:
    """Every function name the C header declares (parsed from include/portblas_amd.h)."""
    import re

    hdr = Path(__file__).resolve().parent.parent / "include" / "portblas_amd.h"
    text = hdr.read_text()
    return sorted(set(re.findall(r"\b(pb_[a-z0-9_]+)\s*\(", text)))


def load() -> ctypes.CDLL:
    """Loads the library (once) and declares every signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m portblas_amd.build` "
            "(there is no fallback implementation)")
    try:  # share torch's HIP runtime when torch is present
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(str(LIB_PATH))
    for name, args in SIGNATURES.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, _i32)
    _lib = L
    return L
