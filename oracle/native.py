"""ctypes access to the CPU checkers built by oracle/Makefile.

TEST INFRASTRUCTURE ONLY (tests/, smoke(), bench.py cpu_baseline).

* ``cic``  -> oracle/libcic_oracle.so : our C restatement of the CIC kernels and
  the bit-exact emulation of the GPU fixed-point paint.
* ``ref``  -> oracle/_ref/libjerald_ref.so : the reference's own
  lib/parops_fast.h CIC drivers at MPI comm size 1 (None when not built).
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_P = C.c_void_p
_I64 = C.c_int64


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _load(path):
    return C.CDLL(path) if os.path.exists(path) else None


_cic = None
_ref = None


def cic_lib():
    global _cic
    if _cic is None:
        _cic = _load(os.path.join(_HERE, "libcic_oracle.so"))
        if _cic is None:
            raise RuntimeError("oracle/libcic_oracle.so not built (run `make -C oracle`)")
        _cic.oracle_paint_f64.argtypes = [_P, _I64, _P, C.c_double, _I64, _I64, _I64, _P, C.c_int]
        _cic.oracle_readout_f64.argtypes = [_P, _I64, _P, _I64, _I64, _I64, C.c_int, _P, _P, _I64, C.c_int]
        _cic.oracle_mesh_units_f64.argtypes = [_P, _I64, _P, _P, _P]
        _cic.oracle_paint_fixed_f32.argtypes = [_P, _I64, C.c_int, _P, _P, _P, _P, C.c_double,
                                                _I64, _I64, _I64, C.c_int, _P]
        _cic.oracle_paint_fixed_f64.argtypes = [_P, _I64, C.c_int, _P, _P, _P, _P, C.c_double,
                                                _I64, _I64, _I64, C.c_int, _P]
    return _cic


def ref_lib():
    """The reference's CIC drivers, or None when oracle/_ref was not built."""
    global _ref
    if _ref is None:
        _ref = _load(os.path.join(_HERE, "_ref", "libjerald_ref.so"))
    return _ref


# ---------------------------------------------------------------- C oracle
def paint_f64(x, mass, nmesh, nthreads=0):
    x = np.ascontiguousarray(x, dtype=np.float64)
    nm = [int(n) for n in nmesh]
    out = np.empty(nm, dtype=np.float64)
    if np.ndim(mass) == 0:
        cic_lib().oracle_paint_f64(_ptr(x), x.shape[0], None, float(mass), *nm, _ptr(out), nthreads)
    else:
        m = np.ascontiguousarray(mass, dtype=np.float64)
        cic_lib().oracle_paint_f64(_ptr(x), x.shape[0], _ptr(m), 0.0, *nm, _ptr(out), nthreads)
    return out


def readout_f64(x, field, deriv_dim=-1, cellinv=(1.0, 1.0, 1.0), nthreads=0):
    x = np.ascontiguousarray(x, dtype=np.float64)
    f = np.ascontiguousarray(field, dtype=np.float64)
    ci = np.ascontiguousarray(cellinv, dtype=np.float64)
    out = np.empty(x.shape[0], dtype=np.float64)
    cic_lib().oracle_readout_f64(_ptr(x), x.shape[0], _ptr(f), *f.shape, int(deriv_dim), _ptr(ci),
                                 _ptr(out), 1, nthreads)
    return out


def readout_grad_f64(x, field, cellinv, nthreads=0):
    return np.stack([readout_f64(x, field, d, cellinv, nthreads) for d in range(3)], axis=-1)


def fixed_point_shift(np_, maxabs, dtype=np.float32):
    """Fixed-point exponent s of the paint (deposits are rint(w m 2^s)), with
    max|m| < 2^e: FP32 s = 30 - e (every deposit fits in 31 bits); FP64
    s = 61 - e - ceil(log2 Np) (the total stays below 2^61)."""
    if maxabs == 0.0 or np_ == 0:
        return 0
    e = math.frexp(maxabs)[1]
    if np.dtype(dtype) == np.float32:
        return 30 - e
    lg = int(np_ - 1).bit_length()
    return 61 - e - lg


def paint_fixed(pos, mass, nmesh, box=None, lattice=None, lattice_step=None, s=None, dtype=np.float32):
    """Bit-exact emulation of the GPU fixed-point paint.  Returns (acc int64
    mesh, s).  ``pos`` physical (box given) or lattice displacements in mesh units."""
    dtype = np.dtype(dtype)
    pos = np.ascontiguousarray(pos, dtype=dtype)
    nm = [int(n) for n in nmesh]
    np_ = pos.shape[0]
    scalar = np.ndim(mass) == 0
    if scalar:
        maxabs = abs(float(dtype.type(mass)))
        m = None
    else:
        m = np.ascontiguousarray(mass, dtype=dtype)
        maxabs = float(np.max(np.abs(m))) if np_ else 0.0
    if s is None:
        s = fixed_point_shift(np_, maxabs, dtype)
    acc = np.empty(nm, dtype=np.int64)
    kind = 0 if lattice is None else 1
    boxa = np.ascontiguousarray(np.broadcast_to(np.asarray(box if box is not None else 1.0, np.float64), (3,)))
    lat = np.ascontiguousarray(lattice if lattice is not None else (1, 1, 1), dtype=np.int64)
    lst = np.ascontiguousarray(lattice_step if lattice_step is not None else (1.0, 1.0, 1.0), dtype=np.float64)
    fn = cic_lib().oracle_paint_fixed_f32 if dtype == np.float32 else cic_lib().oracle_paint_fixed_f64
    fn(_ptr(pos), np_, kind, _ptr(boxa), _ptr(lat), _ptr(lst), _ptr(m),
       float(dtype.type(mass)) if scalar else 0.0, *nm, int(s), _ptr(acc))
    return acc, s


def fixed_to_float(acc, s, dtype=np.float32, scale=1.0):
    """acc * 2^-s * scale, computed in double and rounded once to ``dtype``."""
    return (acc.astype(np.float64) * math.ldexp(1.0, -s) * scale).astype(dtype)


# ---------------------------------------------------------- reference CIC
def _ref_call(name, dtype):
    lib = ref_lib()
    if lib is None:
        raise RuntimeError("oracle/_ref not built")
    fn = getattr(lib, f"{name}_{'f32' if np.dtype(dtype) == np.float32 else 'f64'}")
    return fn


def ref_paint(x, mass, nmesh, dtype=np.float64):
    x = np.ascontiguousarray(x, dtype=dtype)
    m = np.ascontiguousarray(np.broadcast_to(np.asarray(mass, dtype), (x.shape[0],)))
    n = np.ascontiguousarray(nmesh, dtype=np.int64)
    out = np.empty([int(v) for v in nmesh], dtype=dtype)
    fn = _ref_call("ref_paint", dtype)
    fn.argtypes = [_P, _I64, _P, _P, _P]
    fn(_ptr(x), x.shape[0], _ptr(m), _ptr(n), _ptr(out))
    return out


def ref_readout(x, field, box, vjpdim=-1, dtype=np.float64):
    x = np.ascontiguousarray(x, dtype=dtype)
    f = np.ascontiguousarray(field, dtype=dtype)
    n = np.ascontiguousarray(f.shape, dtype=np.int64)
    b = np.ascontiguousarray(np.broadcast_to(np.asarray(box, dtype), (3,)))
    out = np.empty(x.shape[0], dtype=dtype)
    scratch = np.empty(f.shape, dtype=dtype)
    fn = _ref_call("ref_readout", dtype)
    fn.argtypes = [_P, _I64, _P, _P, _P, C.c_int, _P, _P]
    fn(_ptr(x), x.shape[0], _ptr(f), _ptr(n), _ptr(b), int(vjpdim), _ptr(out), _ptr(scratch))
    return out


def ref_readout3(x, field, box, vjp, dtype=np.float64):
    """vjp=False: field is (3, N0, N1, N2); vjp=True: field is (N0, N1, N2)."""
    x = np.ascontiguousarray(x, dtype=dtype)
    f = np.ascontiguousarray(field, dtype=dtype)
    shp = f.shape if vjp else f.shape[1:]
    n = np.ascontiguousarray(shp, dtype=np.int64)
    b = np.ascontiguousarray(np.broadcast_to(np.asarray(box, dtype), (3,)))
    out = np.empty((x.shape[0], 3), dtype=dtype)
    scratch = np.empty(shp, dtype=dtype)
    fn = _ref_call("ref_readout3", dtype)
    fn.argtypes = [_P, _I64, _P, _P, _P, C.c_int, _P, _P]
    fn(_ptr(x), x.shape[0], _ptr(f), _ptr(n), _ptr(b), int(bool(vjp)), _ptr(out), _ptr(scratch))
    return out


def ref_paint3(x, mass3, nmesh, dtype=np.float64):
    x = np.ascontiguousarray(x, dtype=dtype)
    m = np.ascontiguousarray(mass3, dtype=dtype)
    n = np.ascontiguousarray(nmesh, dtype=np.int64)
    out = np.empty([3] + [int(v) for v in nmesh], dtype=dtype)
    fn = _ref_call("ref_paint3", dtype)
    fn.argtypes = [_P, _I64, _P, _P, _P]
    fn(_ptr(x), x.shape[0], _ptr(m), _ptr(n), _ptr(out))
    return out
