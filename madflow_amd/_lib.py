"""Loader for the in-tree HIP library ``libmadflow_amd.so`` (C ABI in include/madflow_amd.h).

The product path has no CPU fallback: if the library is missing or no GPU is
visible, the calls raise.  ``madflow_amd.build.build()`` compiles the library.
"""

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libmadflow_amd.so")

_lib = None

c_i64 = ctypes.c_int64
c_int = ctypes.c_int
c_dbl = ctypes.c_double
c_vp = ctypes.c_void_p
c_u32 = ctypes.c_uint32
c_u64 = ctypes.c_uint64


class MfxError(RuntimeError):
    pass


_SIGNATURES = {
    "mfx_version": ([], c_int),
    "mfx_last_error": ([], ctypes.c_char_p),
    "mfx_process_count": ([], c_int),
    "mfx_process_name": ([c_int], ctypes.c_char_p),
    "mfx_process_index": ([ctypes.c_char_p], c_int),
    "mfx_process_info": ([c_int, c_vp], c_int),
    "mfx_smatrix": ([c_int, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_int, c_int, c_vp, c_i64, c_vp], c_int),
    "mfx_smatrix_workspace_size": ([c_int, c_i64, c_int, c_int], c_i64),
    "mfx_smatrix_noalloc": ([c_int, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_int, c_int, c_vp], c_int),
    "mfx_smatrix_noalloc_scratch_bytes": ([], c_i64),
    "mfx_matrix": ([c_int, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_int, c_int, c_vp, c_i64, c_vp], c_int),
    "mfx_wavefunction": ([c_int, c_vp, c_i64, c_dbl, c_int, c_int, c_vp, c_int, c_int, c_vp], c_int),
    "mfx_aloha": ([c_int, c_vp, c_int, c_i64, c_vp, c_dbl, c_dbl, c_vp, c_int, c_vp], c_int),
    "mfx_rambo": ([c_vp, c_i64, c_int, c_int, c_vp, ctypes.c_double, c_vp, c_vp, c_vp, c_vp, c_vp], c_int),
    "mfx_philox_uniforms": ([c_u64, c_i64, c_int, c_u32, c_u32, c_vp, c_vp], c_int),
    "mfx_philox_uniforms_f32": ([c_u64, c_i64, c_int, c_u32, c_u32, c_vp, c_vp], c_int),
    "mfx_integrand_workspace_size": ([c_i64, c_int, c_int], c_i64),
    "mfx_unweight_workspace_size": ([c_i64], c_i64),
    "mfx_unweight": ([c_vp, c_i64, c_dbl, c_u32, c_u32, c_u64, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp], c_int),
}


def load(path=None):
    """Load the library once (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise MfxError(
            "madflow_amd: %s not found — build it first (python -m madflow_amd.build)" % path)
    lib = ctypes.CDLL(path)
    for name, (args, res) in _SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def exported_symbols():
    """Names declared in include/madflow_amd.h (used by the CPU symbol test)."""
    import re
    hdr = os.path.join(os.path.dirname(PKG), "include", "madflow_amd.h")
    text = open(hdr).read()
    return sorted(set(re.findall(r"\b(mfx_[a-z0-9_]+)\s*\(", text)))


def check(rc):
    """Raise MfxError with mfx_last_error() for a non-zero return code."""
    if rc != 0:
        msg = load().mfx_last_error()
        raise MfxError("madflow_amd call failed (%d): %s" % (rc, msg.decode() if msg else ""))
    return rc


def stream_ptr(device=None):
    """HIP stream handle of torch's current stream (the kernels are enqueued on it)."""
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def require_device_tensor(t, dtype=None, name="tensor"):
    import torch
    if not isinstance(t, torch.Tensor):
        raise TypeError("%s must be a torch tensor on the GPU" % name)
    if not t.is_cuda:
        raise MfxError("%s must live on the GPU (madflow_amd has no CPU path)" % name)
    if dtype is not None and t.dtype != dtype:
        raise TypeError("%s must have dtype %s, got %s" % (name, dtype, t.dtype))
    return t.contiguous()
