"""ctypes binding of libpfhip.so (include/pfhip.h).

torch is imported first so that the library binds to torch's HIP runtime and
rocFFT already loaded in the process.  There is no fallback: if the library
is missing the import of an operator fails with a clear error.
"""

from __future__ import annotations

import ctypes
import os

import torch  # noqa: F401  (loads the HIP runtime the library must share)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpfhip.so")

PF_NORM_NONE = 0
PF_NORM_ORTHO = 1
PF_NORM_ADJOINT = 2
PF_TOEP_RMAX = 128  # include/pfhip.h: largest Gram kernel block nseg * ncoeff
PF_NOALLOC_SCRATCH_BYTES = 64 << 20  # include/pfhip.h: pf_nufft_noalloc's fixed workspace

_ERRORS = {1: "invalid argument", 2: "HIP error", 3: "rocFFT error", 4: "allocation failed", 5: "unsupported"}


class PFError(RuntimeError):
    """Error raised by the native NUFFT engine."""


class BatchDesc(ctypes.Structure):
    _fields_ = [
        ("batch", ctypes.c_int),
        ("nseg", ctypes.c_int),
        ("ncoeff", ctypes.c_int),
        ("basis", ctypes.c_void_p),
        ("b0_temporal", ctypes.c_void_p),
        ("b0_spatial", ctypes.c_void_p),
        ("nt", ctypes.c_int),
    ]


class PlanInfo(ctypes.Structure):
    _fields_ = [
        ("ndim", ctypes.c_int),
        ("shape", ctypes.c_int * 3),
        ("os_shape", ctypes.c_int * 3),
        ("width", ctypes.c_int * 3),
        ("beta", ctypes.c_float * 3),
        ("scale", ctypes.c_double),
        ("nframes", ctypes.c_int),
        ("npts", ctypes.c_longlong),
        ("tile", ctypes.c_int * 3),
        ("spread_entries", ctypes.c_longlong),
        ("spread_items", ctypes.c_int),
        ("spread_slots", ctypes.c_int),
    ]


_VP = ctypes.c_void_p
_I = ctypes.c_int
_SIG = {
    "pf_last_error": (ctypes.c_char_p, []),
    "pf_version": (_I, []),
    "pf_plan_create": (
        _I,
        [_VP, _I, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(ctypes.c_float), _I, ctypes.c_longlong, _VP,
         ctypes.POINTER(_VP)],
    ),
    "pf_plan_create_tiled": (
        _I,
        [_VP, _I, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(ctypes.c_float), _I, ctypes.c_longlong,
         ctypes.POINTER(_I), _I, _VP, ctypes.POINTER(_VP)],
    ),
    "pf_spread_tiles": (_I, [_I, ctypes.POINTER(_I), _I, ctypes.POINTER(_I)]),
    "pf_plan_create_ragged": (
        _I,
        [_VP, _I, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(ctypes.c_float), _I,
         ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(_I), _I, _VP, ctypes.POINTER(_VP)],
    ),
    "pf_plan_destroy": (_I, [_VP]),
    "pf_plan_get_info": (_I, [_VP, ctypes.POINTER(PlanInfo)]),
    "pf_plan_get_permutation": (_I, [_VP, _VP, _VP]),
    "pf_workspace_size": (_I, [_VP, ctypes.POINTER(BatchDesc), _I, ctypes.POINTER(ctypes.c_size_t)]),
    "pf_nufft": (_I, [_VP, ctypes.POINTER(BatchDesc), _VP, _VP, _VP, _I, _VP, ctypes.c_size_t, _VP]),
    "pf_nufft_adj": (_I, [_VP, ctypes.POINTER(BatchDesc), _VP, _VP, _VP, _I, _VP, ctypes.c_size_t, _VP]),
    "pf_nufft_noalloc": (_I, [_VP, ctypes.POINTER(BatchDesc), _I, _VP, _VP, _VP, _I, _VP]),
    "pf_describe_launch": (_I, [_VP, ctypes.POINTER(BatchDesc), ctypes.c_char_p, ctypes.c_size_t]),
    "pf_grid_size": (_I, [_VP, ctypes.POINTER(BatchDesc), ctypes.POINTER(ctypes.c_size_t)]),
    "pf_deapod_pad": (_I, [_VP, ctypes.POINTER(BatchDesc), _VP, _VP, _VP]),
    "pf_fft": (_I, [_VP, ctypes.POINTER(BatchDesc), _VP, _I, _VP]),
    "pf_interp": (_I, [_VP, ctypes.POINTER(BatchDesc), _VP, _VP, _VP]),
    "pf_spread": (_I, [_VP, ctypes.POINTER(BatchDesc), _VP, _VP, _VP, _VP]),
    "pf_spread_workspace_size": (_I, [_VP, ctypes.POINTER(BatchDesc), ctypes.POINTER(ctypes.c_size_t)]),
    "pf_crop_deapod": (_I, [_VP, ctypes.POINTER(BatchDesc), _VP, _VP, ctypes.c_float, _VP]),
    "pf_workspace_offsets": (
        _I,
        [_VP, ctypes.POINTER(BatchDesc)] + [ctypes.POINTER(ctypes.c_size_t)] * 4,
    ),
    "pf_pad_fft": (_I, [_VP, ctypes.POINTER(BatchDesc), _VP, _VP, _VP, _VP]),
    "pf_interp_grad": (_I, [_VP, ctypes.POINTER(BatchDesc), _VP, _VP, ctypes.POINTER(ctypes.c_float), _I, _VP, _VP]),
    "pf_toeplitz_create":(_I, [_VP, _I, ctypes.POINTER(_I), _I, _I, _VP, ctypes.POINTER(_VP)]),
    "pf_toeplitz_destroy": (_I, [_VP]),
    "pf_toeplitz_workspace_size": (_I, [_VP, _I, ctypes.POINTER(ctypes.c_size_t)]),
    "pf_toeplitz_apply": (_I, [_VP, _I, _VP, _VP, _VP, _VP, ctypes.c_size_t, _VP]),
    "pf_gram_create": (_I, [_VP, _I, ctypes.POINTER(_I), _I, _I, _I, _VP, ctypes.POINTER(_VP)]),
    "pf_cart_create": (_I, [_VP, _I, _I, ctypes.POINTER(_I), _VP, ctypes.POINTER(_VP)]),
    "pf_cart_destroy": (_I, [_VP]),
    "pf_cart_workspace_size": (_I, [_VP, ctypes.POINTER(BatchDesc), _I, ctypes.POINTER(ctypes.c_size_t)]),
    "pf_cart_fft": (_I, [_VP, ctypes.POINTER(BatchDesc), _I, _VP, _VP, _I, _VP, ctypes.c_size_t, _VP]),
    "pf_cart_ifft": (_I, [_VP, ctypes.POINTER(BatchDesc), _I, _VP, _VP, _I, _VP, ctypes.c_size_t, _VP]),
    "pf_grog_index": (_I, [_VP, _I, ctypes.POINTER(_I), ctypes.c_longlong, _I, _VP, _VP, _VP, _VP]),
    "pf_grog_apply": (_I, [_VP, _VP, _I, _I, ctypes.c_longlong, _I, _I, _VP, _I, _VP, _VP]),
    "pf_ifft_crop": (_I, [_VP, ctypes.POINTER(BatchDesc), _VP, _VP, ctypes.c_float, _VP, _VP]),
    "pf_sort_samples": (_I, [_VP, ctypes.POINTER(BatchDesc), _VP, _VP, _VP, _VP]),
    "pf_unsort_samples": (_I, [_VP, ctypes.POINTER(BatchDesc), _VP, _VP, _VP, ctypes.c_float, _VP]),
    "pf_voronoi_area_2d": (_I, [_VP, _VP, _I, _I, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_longlong,
                                _VP, _VP, ctypes.c_longlong, _VP, _I, _VP, _VP, _VP]),
}

EXPORTED = tuple(_SIG)

_lib = None


def lib():
    """Load libpfhip.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python pyfourier_amd/csrc/build.py` "
                "(pyfourier_amd has no CPU or PyTorch fallback path)"
            )
        h = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIG.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        msg = lib().pf_last_error().decode(errors="replace")
        raise PFError(f"pyfourier_amd native error {rc} ({_ERRORS.get(rc, '?')}): {msg}")


def ptr(t: "torch.Tensor | None"):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def stream_of(device: "torch.device"):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
