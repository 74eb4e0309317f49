"""ctypes binding of libpyr3hip.so (the C ABI in include/pyr3hip.h).

The library is the only compute path: there is no CPU fallback. Loading fails
with an ImportError naming the missing file when the in-tree build is absent,
so a host without the HIP extension stops instead of silently degrading.
"""
from __future__ import annotations

import ctypes as C
import os
import sys
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libpyr3hip.so"

# Every entry point the header declares: name -> (restype, argtypes).
_P = C.c_void_p
_I = C.c_int
_SZ = C.c_size_t
_FP = C.POINTER(C.c_float)
_UP = C.POINTER(C.c_uint32)
_BP = C.POINTER(C.c_uint8)
SIGNATURES = {
    "p3_create": (_I, [_I, _I, _I, _FP, _I, C.POINTER(_P)]),
    "p3_resize": (_I, [_P, _I, _I, _I, _FP, _I]),
    "p3_destroy": (None, [_P]),
    "p3_reset": (_I, [_P, _FP]),
    "p3_iterate": (_I, [_P, _FP, _I, _UP, _I, _UP]),
    "p3_iterate_multi": (_I, [_P, _I, C.POINTER(_FP), _I, C.POINTER(_UP), _I, C.POINTER(_UP)]),
    "p3_iterate_states": (_I, [_P, _FP, _I, _UP, _I, _UP, _UP]),
    "p3_iterate_noalloc": (_I, [_FP, _I, _UP, _I, _UP, _FP, _P, _SZ, _P]),
    "p3_present": (_I, [_P, _UP, _UP]),
    "p3_read_histogram": (_I, [_P, _UP, _SZ]),
    "p3_read_index": (_I, [_P, _UP, _SZ]),
    "p3_set_curves": (_I, [_P, _FP, _SZ]),
    "p3_read_linear": (_I, [_P, C.c_double, C.c_double, _FP, _SZ]),
    "p3_probe_atomics": (_I, [_P, C.c_uint32, C.c_uint32, C.POINTER(C.c_double),
                              C.POINTER(C.c_int)]),
    "p3_read_filtered": (_I, [_P, _FP, _SZ]),
    "p3_read_image": (_I, [_P, _BP, _SZ]),
    "p3_read_image_f32": (_I, [_P, _FP, _SZ]),
    "p3_write_histogram": (_I, [_P, _UP, _SZ]),
    "p3_set_trace": (_I, [_P, _I, _I]),
    "p3_read_trace": (_I, [_P, _FP, _SZ]),
    "p3_set_image_f32": (_I, [_P, _I]),
    "p3_set_timing": (_I, [_P, _I]),
    "p3_read_timings": (_I, [_P, _FP]),
    "p3_denoiser_param_count": (_SZ, []),
    "p3_denoiser_create": (_I, [_FP, _SZ, C.POINTER(_P)]),
    "p3_denoiser_destroy": (_I, [_P]),
    "p3_denoise": (_I, [_P, _P]),
    "p3_sync": (_I, [_P]),
    "p3_dims": (_I, [_P, C.POINTER(_I)]),
    "p3_chaos_prepare": (_I, [_FP, _I, _I, _I]),
    "p3_chaos_ready": (_I, [_FP, _I, _I, _I]),
    "p3_chaos_source": (_I, [_FP, _I, _I, _I, C.c_char_p, _SZ, C.POINTER(_SZ)]),
    "p3_variation_supported": (_I, [_I]),
    "p3_last_error": (C.c_char_p, [_P]),
    "p3_batch_create": (_I, [C.POINTER(_P)]),
    "p3_batch_destroy": (None, [_P]),
    "p3_batch_add": (_I, [_P, _I, _I, _I, _FP, _I, _FP, _I, C.POINTER(_FP), _I, C.POINTER(_UP),
                          _I, C.POINTER(_UP), _UP, _UP]),
    "p3_batch_commit": (_I, [_P, _I, _I]),
    "p3_batch_launch": (_I, [_P]),
    "p3_batch_sync": (_I, [_P]),
    "p3_batch_elapsed": (_I, [_P, _FP]),
    "p3_batch_frames": (_I, [_P, C.POINTER(_I)]),
    "p3_batch_stats": (_I, [_P, _UP]),
    "p3_batch_read_image": (_I, [_P, _I, _BP, _SZ]),
    "p3_batch_read_histogram": (_I, [_P, _I, _UP, _SZ]),
    "p3_batch_last_error": (C.c_char_p, [_P]),
    "p3_union_prepare": (_I, [_I, C.POINTER(_FP), C.POINTER(_I), C.POINTER(_I), C.POINTER(_I)]),
    # training kernels of the learned denoiser (include/pyr3nn.h)
    "p3nn_gemm": (_I, [_P, _I, _I, _I, _I, _P]),
    "p3nn_im2col": (_I, [_P, _P, _P, _P, _I, _P]),
    "p3nn_col2im": (_I, [_P, _P, _I, _P, _P, _P]),
    "p3nn_relu_bwd": (_I, [_P, _P, _P, _SZ, _P]),
    "p3nn_bias_grad": (_I, [_P, _SZ, _I, _P, _P]),
    "p3nn_features": (_I, [_P, _P, C.c_float, _SZ, _P, _P]),
    "p3nn_head": (_I, [_P, _P, _P, _P, _SZ, _P, _P, _P, _P, _P]),
    "p3nn_adam": (_I, [_P, _P, _P, _P, _SZ, C.c_float, C.c_float, C.c_float, C.c_float, _I, _P,
                       _P]),
    "p3nn_cast_bf16": (_I, [_P, _P, _SZ, _P]),
}

_lib = None
# Whether torch's modules were loaded when the library was. torch ships its
# own HIP and HSA runtimes (torch/lib); libpyr3hip.so binds to whichever
# libamdhip64.so.7 the process already holds. Loaded after torch, the process
# has one runtime; loaded first, it takes /opt/rocm's, and torch's own copy,
# loaded later, cannot open the GPU ("No HIP GPUs are available").
loaded_after_torch = False


def load() -> C.CDLL:
    """Load the in-tree library once; raise ImportError if it was never built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"pyr3_amd: {LIB_PATH} is missing; build it with `python -m pyr3_amd.build` "
            "(there is no CPU fallback for the render path)")
    # Compiled per-shape chaos kernels persist in the package's kcache/ unless
    # P3_KERNEL_CACHE names another directory ("" turns the cache off).
    os.environ.setdefault("P3_KERNEL_CACHE", str(LIB_PATH.parent / "kcache"))
    global loaded_after_torch
    loaded_after_torch = "torch" in sys.modules
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class P3Error(RuntimeError):
    """A non-zero status from the C ABI, carrying p3_last_error()'s text."""

    def __init__(self, code: int, message: str):
        super().__init__(f"pyr3hip error {code}: {message}")
        self.code = code


def check(code: int, handle=None) -> None:
    if code != 0:
        msg = load().p3_last_error(handle)
        raise P3Error(code, msg.decode() if msg else "")
