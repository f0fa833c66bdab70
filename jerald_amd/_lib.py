"""ctypes binding of libldl.so (include/ldl.h) plus torch plumbing.

The library is loaded from the package directory; a missing library or a
missing GPU raises instead of falling back to any other implementation.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libldl.so")

LDL_POS_PHYSICAL = 0
LDL_POS_LATTICE = 1
LDL_PAINT_FIXED = 0
LDL_PAINT_FLOAT_ATOMIC = 1
LDL_PAINT_WS_ZEROED = 2
LDL_PAINT_DIRECT = 4
LDL_PAINT_BINS_VALID = 8


class LdlError(RuntimeError):
    """A libldl entry point returned a non-zero status."""


class PosDesc(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("reserved", C.c_int32),
        ("box", C.c_double * 3),
        ("lattice", C.c_int64 * 3),
        ("lattice_step", C.c_double * 3),
    ]


_P = C.c_void_p
_I64 = C.c_int64
_I32 = C.c_int32
_D = C.c_double
_SZ = C.c_size_t
_I64x3 = C.POINTER(C.c_int64)
_Dp = C.POINTER(C.c_double)
_PD = C.POINTER(PosDesc)

# name -> argtypes (restype int unless listed in _RESTYPE)
_SIGS = {
    "ldl_last_error": [],
    "ldl_version": [],
    "ldl_buildplan": [_I64x3, C.POINTER(C.c_int32)],
    "ldl_paint_workspace_bytes": [_I64x3, _I32, _I64, _I32],
    "ldl_fft_workspace_bytes": [_I64x3, _I32],
    "ldl_sm_workspace_bytes": [_I64x3, _I32, _I64, _I32],
    "ldl_sm_bins_bytes": [_I64x3, _I64],
    "ldl_sm_spectrum_pitch": [_I64x3, _I32],
    "ldl_timer_enable": [_I32],
    "ldl_timer_read": [C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_int64)],
}
for _t in ("f32", "f64"):
    _SIGS.update({
        f"ldl_ppaint_{_t}": [_P, _PD, _I64, _P, _D, _I64x3, _D, _I32, _P, _P, _SZ, _P],
        f"ldl_ppaint3D_{_t}": [_P, _PD, _I64, _P, _I64x3, _D, _I32, _P, _P, _SZ, _P],
        f"ldl_preadout_{_t}": [_P, _PD, _I64, _P, _I64x3, _I32, _P, _P],
        f"ldl_preadout3D_{_t}": [_P, _PD, _I64, _P, _I64x3, _I32, _P, _P, _P],
        f"ldl_readout_grad3_{_t}": [_P, _PD, _I64, _P, _P, _I64x3, _P, _P],
        f"ldl_pfft_{_t}": [_P, _P, _I64x3, _D, _P],
        f"ldl_pifft_{_t}": [_P, _P, _I64x3, _D, _P, _SZ, _P],
        f"ldl_sm_layer_fwd_{_t}": [_P, _PD, _I64, _I64x3, _Dp, _D, _I32, _P, _P, _P, _P, _P, _P, _SZ, _P],
        f"ldl_sm_layer_bwd_{_t}": [_P, _PD, _I64, _I64x3, _Dp, _D, _I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _SZ, _P],
        f"ldl_sm_loss_{_t}": [_P, _P, _P, _P, _P, _I64x3, _Dp, _Dp, _D, _I32, _I32, _D, _P, _P, _P, _SZ, _P],
        f"ldl_dm_layer_fwd_{_t}": [_P, _PD, _I64, _I64x3, _Dp, _D, _I32, _P, _P, _P, _P, _P, _P, _SZ, _P],
        f"ldl_dm_layer_bwd_{_t}": [_P, _PD, _I64, _I64x3, _Dp, _D, _I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _SZ, _P],
        f"ldl_dm_loss_{_t}": [_P, _P, _P, _P, _P, _I64x3, _Dp, _D, _I32, _I32, _D, _P, _P, _P, _SZ, _P],
        f"ldl_hi_layer_fwd_{_t}": [_P, _PD, _I64, _I64x3, _Dp, _D, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _SZ, _P],
        f"ldl_hi_layer_bwd_{_t}": [_P, _PD, _I64, _I64x3, _Dp, _D, _I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                   _P, _SZ, _P],
        f"ldl_hi_loss_{_t}": [_P, _P, _P, _P, _P, _P, _I64x3, _Dp, _Dp, _D, _I32, _I32, _D, _P, _P, _P, _P, _SZ, _P],
        f"ldl_hi_depletion_fwd_{_t}": [_P, _P, _I64x3, _Dp, _Dp, _P, _P, _P, _SZ, _P],
        f"ldl_hi_depletion_bwd_{_t}": [_P, _P, _P, _P, _I64x3, _Dp, _Dp, _I32, _P, _P, _SZ, _P],
        f"ldl_paint_vjp_pos_{_t}": [_P, _PD, _I64, _P, _I64x3, _D, _P, _P, _P],
        f"ldl_spectrum_zweight_{_t}": [_P, _P, _I64x3, _I32, _D, _P],
        f"ldl_ppaint_noalloc_{_t}": [_P, _PD, _I64, _P, _D, _I64x3, _D, _I32, _P, _P],
        f"ldl_sm_response_{_t}": [_P, _I64x3, _Dp, _P, _P],
        f"ldl_hi_response_{_t}": [_P, _P, _I64x3, _Dp, _P, _P],
    })
_RESTYPE = {
    "ldl_last_error": C.c_char_p,
    "ldl_paint_workspace_bytes": C.c_size_t,
    "ldl_fft_workspace_bytes": C.c_size_t,
    "ldl_sm_workspace_bytes": C.c_size_t,
    "ldl_sm_bins_bytes": C.c_size_t,
    "ldl_sm_spectrum_pitch": C.c_int64,
}

EXPORTED = sorted(_SIGS)

_lib = None
_lock = threading.Lock()


def load(path: str = LIB_PATH):
    """Load libldl.so once; raises if it has not been built."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(path):
                raise LdlError(f"{path} is missing: build it with `python -m jerald_amd.build`")
            lib = C.CDLL(path)
            for name, args in _SIGS.items():
                fn = getattr(lib, name)
                fn.argtypes = args
                fn.restype = _RESTYPE.get(name, C.c_int)
            _lib = lib
    return _lib


def call(name: str, *args):
    """Call an int-returning entry point and raise LdlError on failure."""
    rc = getattr(load(), name)(*args)
    if rc != 0:
        msg = load().ldl_last_error().decode(errors="replace")
        raise LdlError(f"{name} failed ({rc}): {msg}")


def timer_enable(on: bool):
    call("ldl_timer_enable", int(bool(on)))


def timer_read(name: str):
    """(total device ms, launch count) of an instrumented kernel."""
    t = C.c_double(0.0)
    n = C.c_int64(0)
    call("ldl_timer_read", name.encode(), C.byref(t), C.byref(n))
    return t.value, n.value


def sfx(dtype: torch.dtype) -> str:
    if dtype == torch.float32:
        return "f32"
    if dtype == torch.float64:
        return "f64"
    raise NotImplementedError(f"Unsupported dtype {dtype}")


def i64x3(v) -> C.Array:
    a = (C.c_int64 * 3)(*[int(x) for x in v])
    return a


def dvec(v) -> C.Array:
    return (C.c_double * len(v))(*[float(x) for x in v])


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    return C.c_void_p(t.data_ptr())


def stream_ptr(device) -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def pos_desc(box, kind=LDL_POS_PHYSICAL, lattice=(0, 0, 0), lattice_step=(0.0, 0.0, 0.0)) -> PosDesc:
    d = PosDesc()
    d.kind = kind
    d.reserved = 0
    for i in range(3):
        d.box[i] = float(box[i])
        d.lattice[i] = int(lattice[i])
        d.lattice_step[i] = float(lattice_step[i])
    return d


def require_gpu(t: torch.Tensor, what: str = "tensor"):
    if not t.is_cuda:
        raise LdlError(f"{what} must be a HIP device tensor (got {t.device}); libldl runs on the GPU only")


# ------------------------------------------------------------ workspaces
class _Workspaces:
    """Per-(device, kind, bytes) scratch, allocated zeroed once.  Every
    fixed-point paint leaves its accumulator zero, so paints may pass
    LDL_PAINT_WS_ZEROED after the first use."""

    def __init__(self):
        self._ws = {}

    def get(self, device, kind: str, nbytes: int):
        key = (str(device), kind)
        t = self._ws.get(key)
        if t is None or t.numel() < nbytes:
            t = torch.zeros(max(nbytes, 256), dtype=torch.uint8, device=device)
            self._ws[key] = t
        return t

    def clear(self):
        self._ws.clear()


workspaces = _Workspaces()
