"""ctypes binding of libtd3_hip.so, the gfx950 kernels behind the C ABI of include/td3.h.

This replaces the reference's kernel dispatch (kernels/__init__.py:99-114, a cunumpy
``KernelPackage`` per kernel folder): instead of picking a CUDA/Pyccel/NumPy version by
array type, every entry point here is a hand-written HIP kernel launch on the current
torch stream, taking raw device pointers.  There is no CPU fallback: without the built
library, or without a GPU, calls raise.
"""

from __future__ import annotations

import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtd3_hip.so")

TD3_OK = 0
TD3_ERR_CAPACITY = -2
TD3_ERR_ARG = -3
TD3_ERR_HIP = -4

COULOMB, LOG = 0, 1
FP64, MIXED = 0, 1

_P = ctypes.c_void_p
_I = ctypes.c_int64
_D = ctypes.c_double
_C = ctypes.c_int
_Z = ctypes.c_size_t

# name -> (argtypes, restype); the order of include/td3.h.
SIGNATURES: dict[str, tuple[list, object]] = {
    "td3_abi_version": ([], _C),
    "td3_last_error": ([], ctypes.c_char_p),
    "td3_direct_workspace_bytes": ([_I, _I], _Z),
    "td3_direct": ([_C, _C, _P, _I, _P, _P, _I, _P, _D, _P, _P, _P, _P, _Z, _P], _C),
    "td3_direct_noalloc_scratch_bytes": ([], _Z),
    "td3_direct_noalloc": ([_C, _C, _P, _I, _P, _P, _I, _P, _D, _P, _P, _P, _P], _C),
    "td3_tree_workspace_bytes": ([_I, _I], _Z),
    "td3_tree_build": ([_P, _I, _D, _D, _D, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                        _P, _Z, _P, _P, _P], _C),
    "td3_forest_build": ([_P, _I, _I, _P, _P, _P, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                          _P, _P, _Z, _P, _P, _P], _C),
    "td3_forest_build_nosync": ([_P, _I, _I, _P, _P, _P, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P,
                                 _P, _P, _P, _P, _P, _Z, _P, _P], _C),
    "td3_permute": ([_P, _I, _I, _P, _I, _P, _P], _C),
    "td3_moments": ([_C, _I, _P, _P, _I, _P, _P, _I, _P, _I, _P, _I, _P, _P, _P, _P, _P, _P, _P],
                    _C),
    "td3_coulomb_moments_nosync": ([_P, _P, _I, _P, _P, _I, _P, _I, _P, _P, _P, _P, _P], _C),
    "td3_target_order_workspace_bytes": ([_I], _Z),
    "td3_walk_waves": ([_C], _C),
    "td3_target_order": ([_P, _I, _D, _D, _D, _P, _P, _Z, _P], _C),
    "td3_walk": ([_C, _C, _P, _I, _P, _P, _P, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I,
                  _I, _P, _D, _D, _C, _P, _P, _P, _P], _C),
    "td3_walk_counts": ([_C, _C, _P, _I, _P, _P, _P, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P, _P,
                         _P, _I, _I, _P, _D, _D, _C, _P, _P, _P, _P, _P], _C),
    "td3_traverse_workspace_bytes": ([_I], _Z),
    "td3_traverse": ([_P, _P, _P, _I, _I, _D, _I, _I, _I, _P, _P, _P, _Z, _P, _P], _C),
    "td3_lists_workspace_bytes": ([_I, _I, _C], _Z),
    "td3_lists": ([_P, _I, _C, _I, _I, _P, _P, _P, _P, _Z, _P], _C),
    "td3_scale_moments": ([_C, _P, _P, _P, _I, _I, _P, _P, _P], _C),
    "td3_m2l_workspace_bytes": ([_I, _I, _C], _Z),
    "td3_m2l": ([_C, _I, _P, _P, _P, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _Z, _P], _C),
    "td3_m2l_record_bytes": ([_I], _Z),
    "td3_m2l_records": ([_I, _P, _P, _P, _P, _I, _P, _P], _C),
    "td3_m2l_f32": ([_I, _P, _P, _I, _I, _P, _P, _P, _P, _P, _P, _P], _C),
    "td3_l2l": ([_I, _P, _P, _P, _P, _I, _P, _I, _P, _P, _P], _C),
    "td3_unscale_locals": ([_P, _I, _I, _P, _P, _P], _C),
    "td3_fmm_eval": ([_C, _C, _I, _P, _P, _P, _I, _P, _P, _P, _I, _P, _P, _P, _P, _P, _P, _D, _C,
                      _P, _P, _P, _P], _C),
    "td3_fmm_near": ([_C, _C, _P, _P, _I, _I, _I, _P, _P, _P, _P, _P, _D, _C, _P, _P, _P], _C),
    "td3_fmm_far": ([_C, _I, _P, _P, _I, _P, _P, _P, _P, _P, _P, _C, _P, _P, _P], _C),
    "td3_pack_sources": ([_P, _P, _I, _P, _P], _C),
    "td3_log_moments_dipole": ([_I, _P, _P, _P, _I, _P, _P, _I, _P, _I, _P, _I, _P, _P, _P, _P],
                               _C),
    "td3_log_fmm_eval_ext": ([_I, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _D, _P, _P,
                              _P, _P, _P], _C),
    "td3_log_fmm_at_eval_ext": ([_I, _P, _P, _P, _I, _P, _P, _P, _I, _P, _P, _P, _P, _P, _I, _P,
                                 _P, _P, _D, _P, _P, _P, _P, _P], _C),
    "td3_leaf_index": ([_P, _I, _P, _P], _C),
    "td3_dual_traverse": ([_P, _P, _P, _I, _I, _P, _P, _P, _I, _I, _D, _I, _I, _I, _P, _P, _P, _Z,
                           _P, _P], _C),
    "td3_fmm_at_eval": ([_C, _C, _I, _P, _P, _P, _I, _P, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P,
                         _P, _I, _P, _P, _D, _C, _P, _P, _P, _P], _C),
    "td3_cart_moments": ([_I, _P, _P, _I, _P, _P, _I, _P, _I, _P, _I, _P, _P, _P], _C),
    "td3_cart_m2l_workspace_bytes": ([_I, _I, _I], _Z),
    "td3_cart_m2l": ([_I, _P, _P, _P, _I, _I, _P, _P, _P, _D, _P, _P, _Z, _P], _C),
    "td3_cart_scale_prime": ([_I, _P, _P, _I, _P, _P], _C),
    "td3_cart_m2l_f32": ([_I, _P, _P, _P, _I, _I, _P, _P, _P, _P, _P, _D, _P, _P, _Z, _P], _C),
    "td3_cart_l2l": ([_I, _P, _P, _P, _I, _P, _I, _P, _P], _C),
    "td3_rows_workspace_bytes": ([_P, _I, _I], _Z),
    "td3_rows": ([_P, _P, _P, _P, _I, _D, _I, _I, _I, _P, _P, _P, _P, _P, _P, _Z, _P, _P, _P],
                 _C),
    "td3_push": ([_C, _I, _P, _P, _P, _P, _P, _P, _P, _D, _D, _P], _C),
    "td3_system_bounds": ([_P, _I, _I, _P, _P, _P, _P], _C),
    "td3_soften_near_workspace_bytes": ([_I], _Z),
    "td3_soften_near": ([_P, _I, _P, _P, _P, _P, _P, _I, _P, _P, _P, _I, _D, _D, _C, _P, _P, _P,
                         _P, _Z, _P], _C),
    "td3_system_energies": ([_P, _P, _P, _P, _I, _I, _P, _P, _P], _C),
    "td3_box_tree_build": ([_P, _I, _D, _D, _D, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                            _P, _P, _Z, _P, _P, _P], _C),
    "td3_box_leaf_order": ([_P, _I, _P, _I, _P, _P, _P], _C),
    "td3_smooth_members_workspace_bytes": ([_I, _I], _Z),
    "td3_smooth_members": ([_P, _P, _I, _P, _P, _P, _I, _P, _P, _I, _D, _D, _D, _D, _P, _P, _P,
                            _P, _P, _P, _P, _P, _P, _P, _P, _Z, _P], _C),
    "td3_smooth_eval": ([_I, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _D, _P, _P, _P, _P], _C),
    "td3_smooth_combine_workspace_bytes": ([_I], _Z),
    "td3_smooth_combine": ([_I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _Z, _P], _C),
    "td3_compact_direct_workspace_bytes": ([_I], _Z),
    "td3_compact_direct": ([_P, _I, _C, _P, _P, _I, _D, _P, _P, _P, _P, _P, _Z, _P], _C),
    "td3_timing_enable": ([_C], _C),
    "td3_timing_get": ([ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                        ctypes.POINTER(ctypes.c_int64)], _C),
}

_LIB = None


class CapacityError(RuntimeError):
    """A caller buffer was too small; ``needed`` holds the sizes the library asked for."""

    def __init__(self, message, needed=None):
        super().__init__(message)
        self.needed = needed


def lib() -> ctypes.CDLL:
    """The loaded library; raises if it is not built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is not built: run `python tree-d-3d_amd/build.py` "
                "(hipcc --offload-arch=gfx950)"
            )
        handle = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.argtypes = args
            fn.restype = res
        if handle.td3_abi_version() != 1:
            raise ImportError("libtd3_hip.so ABI version mismatch")
        _LIB = handle
    return _LIB


def last_error() -> str:
    return lib().td3_last_error().decode(errors="replace")


def check(code: int, what: str) -> None:
    if code == TD3_OK:
        return
    message = f"{what}: {last_error()} (code {code})"
    if code == TD3_ERR_CAPACITY:
        raise CapacityError(message)
    if code == TD3_ERR_ARG:
        raise ValueError(message)
    raise RuntimeError(message)


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args), name)


_GPU_SEEN = False
# torch's current raw stream of a device, without building a Stream object (the public
# torch.cuda.current_stream() costs ~10 us of Python per call; a step makes dozens)
_RAW_STREAM = getattr(torch._C, "_cuda_getCurrentRawStream", None)


def device() -> torch.device:
    """The GPU the kernels run on (torch's current device); raises without one."""
    global _GPU_SEEN
    if not _GPU_SEEN:
        if not torch.cuda.is_available():
            raise RuntimeError(
                "tree-d-3d_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
        _GPU_SEEN = True
    return torch.device("cuda", torch.cuda.current_device())


def stream() -> int:
    """torch's current stream on the current device, as a hipStream_t handle."""
    if _RAW_STREAM is not None and _GPU_SEEN:
        return _RAW_STREAM(torch._C._cuda_getDevice())
    return torch.cuda.current_stream().cuda_stream


def ptr(t: torch.Tensor | None) -> int | None:
    if t is None:
        return None
    assert t.is_cuda and t.is_contiguous(), "device arrays must be contiguous GPU tensors"
    return t.data_ptr()


def workspace(nbytes: int) -> torch.Tensor:
    """A scratch buffer from torch's caching allocator (the library never allocates)."""
    return torch.empty(max(int(nbytes), 8), dtype=torch.uint8, device=device())


def timing_enable(on: bool) -> None:
    """Start (clearing) or stop the per-kernel event timing of the library."""
    check(lib().td3_timing_enable(1 if on else 0), "td3_timing_enable")


def timing_get(name: str) -> tuple[float, int]:
    """(summed milliseconds, launches) of the kernels called ``name`` since timing_enable."""
    ms = ctypes.c_double()
    count = ctypes.c_int64()
    check(lib().td3_timing_get(name.encode(), ctypes.byref(ms), ctypes.byref(count)),
          "td3_timing_get")
    return ms.value, count.value
