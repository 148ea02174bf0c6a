"""The HIP backend: loads the C-ABI library (``include/blochparty_amd.h``) with ctypes and launches its entry points
on PyTorch's current stream.

Plays the role of the reference's CUDA backend object (``src/blochparty/engines/gpu/cuda/__init__.py:65-82``:
``name``, ``device``, ``lib()``, ``synchronize()``, ``max_coils``) and of its launcher
(``cuda/rtc.py:116-126``): arguments are raw device pointers, errors come back as ``RuntimeError`` with the
runtime's error string. There is no fallback: without the built library or a ROCm GPU, the engine raises.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

import torch

LIB_PATH = Path(__file__).resolve().parent.parent / "libblochparty_amd.so"
ABI_VERSION = 25
MAX_COILS = 128  # 16 tiles of 8 coils (the forward kernel's LDS budget)


class HbCase(ctypes.Structure):
    """``hb_case`` of the header."""

    _fields_ = [
        ("spins", ctypes.c_void_p),
        ("pos_lo", ctypes.c_void_p),
        ("sens_a", ctypes.c_void_p),
        ("segs", ctypes.c_void_p),
        ("state", ctypes.c_void_p),
        ("maps", ctypes.c_void_p),
        ("n", ctypes.c_int32),
        ("n_pad", ctypes.c_int32),
        ("n_coils", ctypes.c_int32),
        ("n_tiles", ctypes.c_int32),
        ("n_segments", ctypes.c_int32),
        ("n_pulses", ctypes.c_int32),
        ("rf_free", ctypes.c_int32),
        ("sens_scale", ctypes.c_float),
        ("x_scale", ctypes.c_float),
        ("reserved", ctypes.c_int32),
    ]


class HbAdjoint(ctypes.Structure):
    """``hb_adjoint`` of the header."""

    _fields_ = [
        ("ckpt", ctypes.c_void_p),
        ("r8", ctypes.c_void_p),
        ("lam", ctypes.c_void_p),
        ("gs", ctypes.c_void_p),
        ("gc", ctypes.c_void_p),
        ("gseg", ctypes.c_void_p),
        ("h", ctypes.c_void_p),
        ("sens", ctypes.c_void_p),
        ("maps_d", ctypes.c_void_p),
        ("seeds", ctypes.c_void_p),
        ("nu_params", ctypes.c_void_p),
        ("nu_win", ctypes.c_void_p),
        ("nu_grid", ctypes.c_void_p),
        ("nc", ctypes.c_int32),
        ("seg_grads", ctypes.c_int32),
        ("coil_grads", ctypes.c_int32),
        ("k_max", ctypes.c_int32),
    ]


class HbNufft(ctypes.Structure):
    """``hb_nufft`` of the header."""

    _fields_ = [
        ("params", ctypes.c_void_p),
        ("win", ctypes.c_void_p),
        ("part", ctypes.c_void_p),
        ("runs_max", ctypes.c_int32),
        ("cols", ctypes.c_int32),
        ("terms", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


_LIB: ctypes.CDLL | None = None

# (name, argtypes): every entry point returns int (a hipError_t)
_P, _I, _V = ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p
SIGNATURES = {
    "hb_forward": [ctypes.POINTER(HbCase), _I, _I, _P, _I, _P, _V],
    "hb_signal_reduce": [_P, _I, _I, _I, _I, _P, _I, _P, _P, _P, _V],
    "hb_signal_reduce_noalloc": [_P, _I, _I, _I, _I, _P, _I, _P, _P, _V],
    "hb_nufft_grid": [_P, _I, _I, _I, _I, _I, _P, _P, _I, _I, _P, _P, _P, _P, _V],
    "hb_nufft_signal": [_P, _I, _P, _P, _I, _I, _I, _I, _I, _P, _P, _P, _P, _V],
    "hb_nufft_adjoint_grid": [_P, _I, _P, _I, _I, _I, _I, _P, _P, _P, _V],
    "hb_adjoint_interval": [ctypes.POINTER(HbCase), ctypes.POINTER(HbAdjoint), _I, _I, _V],
    "hb_seg_reduce": [_P, _I, _I, _I, _I, _P, _P, _P, _V],
    "hb_pulse_maps": [ctypes.POINTER(HbCase), _P, _P, _I, _P, _I, _P, _V],
    "hb_pulse_map_grads": [ctypes.POINTER(HbCase), _P, _P, _I, _P, _P, _P, _V],
    "hb_pulse_map_dlnb1": [ctypes.POINTER(HbCase), _P, _P, _I, ctypes.c_double, _P, _V],
    "hb_spinor_blocks": [_P, _P, _P, _I, _P, _V],
    "hb_mlp_grad": [_P, _P, _P, _I, _P, _P, _V],
    "hb_mlp_train_step": [_P, _P, _P, _I, _P, _I, ctypes.c_float, _V],
    "hb_mlp_eval": [_P, _P, _I, _P, _V],
    "hb_sur_first_order": [_P, _P, ctypes.c_double, ctypes.c_double, _I, _P, _V],
    "hb_sur_first_quats": [_P, _P, _I, _P, _V],
    "hb_mlp_train_steps": [_P, _P, _P, _I, _I, _P, _I, _I, ctypes.c_float, ctypes.c_float, _V],
    "hb_surrogate_maps": [ctypes.POINTER(HbCase), _P, _P, _P, _P, _V],
    "hb_ifftc_axis": [_P, _P, ctypes.c_int64, _I, ctypes.c_int64, _V],
    "hb_rss": [_P, _I, ctypes.c_int64, _P, _V],
}


def load() -> ctypes.CDLL:
    """The library, loaded once; RuntimeError when it is not built."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"blochparty_amd HIP library not built ({LIB_PATH} missing): run `python -m blochparty_amd.build`"
            )
        lib = ctypes.CDLL(str(LIB_PATH))
        lib.hb_error_string.restype = ctypes.c_char_p
        lib.hb_error_string.argtypes = [ctypes.c_int]
        lib.hb_abi_version.restype = ctypes.c_int
        lib.hb_mlp_workspace.restype = ctypes.c_int64
        lib.hb_mlp_workspace.argtypes = [ctypes.c_int32]
        lib.hb_signal_reduce_workspace.restype = ctypes.c_int64
        lib.hb_signal_reduce_workspace.argtypes = [ctypes.c_int32] * 3
        if lib.hb_abi_version() != ABI_VERSION:
            raise RuntimeError(f"{LIB_PATH} has ABI {lib.hb_abi_version()}, expected {ABI_VERSION}: rebuild it")
        for name, args in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = ctypes.c_int
            fn.argtypes = args
        _LIB = lib
    return _LIB


def ptr(t: torch.Tensor | None) -> int | None:
    if t is None:
        return None
    if not t.is_contiguous():
        raise ValueError("device buffers must be contiguous")
    return t.data_ptr()


class Kernels:
    """``lib.<entry point>(*args)``: calls the C ABI on the current stream, raising RuntimeError on failure."""

    def __getattr__(self, name: str):
        if name.startswith("_"):
            raise AttributeError(name)
        lib = load()
        fn = getattr(lib, "hb_" + name)

        def call(*args):
            stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
            rc = fn(*args, stream)
            if rc != 0:
                raise RuntimeError(f"HIP error {rc} in hb_{name}: {lib.hb_error_string(rc).decode()}")

        return call


class Hip:
    """The backend object of the GPU engine (the reference's ``_Cuda`` role)."""

    name = "hip"
    max_coils = MAX_COILS

    @property
    def device(self) -> torch.device:
        return torch.device("cuda", torch.cuda.current_device())

    def lib(self) -> Kernels:
        return Kernels()

    def synchronize(self) -> None:
        torch.cuda.synchronize()


HIP = Hip()


def available() -> bool:
    """A ROCm GPU is visible to PyTorch and the library is built."""
    try:
        return bool(torch.version.hip) and torch.cuda.is_available() and LIB_PATH.exists()
    except Exception:
        return False
