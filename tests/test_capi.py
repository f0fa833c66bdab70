"""CPU checks of the C ABI: the library loads, exports every entry point
include/ldl.h declares, and rejects bad arguments before touching a GPU."""
import ctypes
import os
import re

import pytest

from jerald_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "ldl.h")).read()
    return sorted(set(re.findall(r"\b(ldl_[a-zA-Z0-9_]+)\s*\(", src)))


def test_header_and_binding_agree():
    assert _declared() == L.EXPORTED


def test_library_exports_every_declared_symbol():
    lib = L.load()
    for name in _declared():
        assert hasattr(lib, name), name
    assert lib.ldl_version() == 1


def test_workspace_queries():
    lib = L.load()
    nm = L.i64x3((64, 64, 64))
    one = lib.ldl_paint_workspace_bytes(nm, 1, 1000, L.LDL_PAINT_FIXED)
    # the tiled paint: sort keys and indices, no accumulator mesh
    assert 2 * 1000 * 4 <= one < 64 ** 3 * 8
    # the direct paint adds one 64-bit accumulator per cell and component
    assert lib.ldl_paint_workspace_bytes(nm, 1, 1000, L.LDL_PAINT_DIRECT) >= one + 64 ** 3 * 8
    assert lib.ldl_paint_workspace_bytes(nm, 3, 1000, L.LDL_PAINT_DIRECT) >= one + 3 * 64 ** 3 * 8
    assert lib.ldl_paint_workspace_bytes(nm, 1, 1000, L.LDL_PAINT_FLOAT_ATOMIC) == 0
    assert lib.ldl_paint_workspace_bytes(nm, 2, 1000, 0) == 0
    assert lib.ldl_paint_workspace_bytes(nm, 1, -1, 0) == 0
    assert lib.ldl_fft_workspace_bytes(nm, 0) == 64 * 64 * 33 * 8
    assert lib.ldl_sm_workspace_bytes(nm, 1, 1000, 0) > lib.ldl_sm_workspace_bytes(nm, 0, 1000, 0) > 0
    assert (lib.ldl_sm_workspace_bytes(nm, 0, 1000, L.LDL_PAINT_DIRECT)
            >= lib.ldl_sm_workspace_bytes(nm, 0, 1000, 0) + 3 * 64 ** 3 * 8)
    # spectrum rows padded to 64 bytes: 16^3 -> 9 complex -> 16 (complex64), 12 (complex128)
    assert lib.ldl_sm_spectrum_pitch(L.i64x3((16, 16, 16)), 0) == 16
    assert lib.ldl_sm_spectrum_pitch(L.i64x3((16, 16, 16)), 1) == 12
    assert lib.ldl_sm_spectrum_pitch(L.i64x3((8, 8, 512)), 0) == 264
    assert lib.ldl_sm_spectrum_pitch(L.i64x3((8, 8, 14)), 0) == 0


@pytest.mark.parametrize("shape", [(14, 8, 8), (8, 8, 7), (8, 8, 15), (2048, 8, 8), (8, 8, 22), (3, 8, 8), (8, 1100, 8)])
def test_buildplan_rejects_unsupported_meshes(shape):
    """FFT axes: [4, 1024], prime factors 2, 3, 5 only, last axis even."""
    with pytest.raises(L.LdlError, match="FFT mesh"):
        L.call("ldl_buildplan", L.i64x3(shape), None)


def test_argument_validation_errors():
    pd = L.pos_desc((1.0, 1.0, 1.0))
    with pytest.raises(L.LdlError, match="np must be >= 0"):
        L.call("ldl_preadout_f32", ctypes.c_void_p(16), pd, -1, ctypes.c_void_p(16), L.i64x3((4, 4, 4)), -1,
               ctypes.c_void_p(16), None)
    with pytest.raises(L.LdlError, match="vjpdim"):
        L.call("ldl_preadout_f32", ctypes.c_void_p(16), pd, 5, ctypes.c_void_p(16), L.i64x3((4, 4, 4)), 3,
               ctypes.c_void_p(16), None)
    bad = L.pos_desc((1.0, -1.0, 1.0))
    with pytest.raises(L.LdlError, match="box sizes"):
        L.call("ldl_preadout_f64", ctypes.c_void_p(16), bad, 5, ctypes.c_void_p(16), L.i64x3((4, 4, 4)), -1,
               ctypes.c_void_p(16), None)
    lat = L.pos_desc((1.0, 1.0, 1.0), L.LDL_POS_LATTICE, (2, 2, 2), (1.0, 1.0, 1.0))
    with pytest.raises(L.LdlError, match="lattice"):
        L.call("ldl_preadout_f32", ctypes.c_void_p(16), lat, 7, ctypes.c_void_p(16), L.i64x3((4, 4, 4)), -1,
               ctypes.c_void_p(16), None)
    with pytest.raises(L.LdlError, match="workspace too small"):
        L.call("ldl_ppaint_f32", ctypes.c_void_p(16), pd, 5, None, 1.0, L.i64x3((4, 4, 4)), 1.0, 0,
               ctypes.c_void_p(16), ctypes.c_void_p(16), 10, None)
    with pytest.raises(L.LdlError, match="p_norm"):
        L.call("ldl_sm_loss_f32", ctypes.c_void_p(16), ctypes.c_void_p(16), None, None, None, L.i64x3((8, 8, 8)),
               L.dvec((1, 1, 1)), L.dvec((1, 1, 0)), 1.0, 3, 0, 0.0, ctypes.c_void_p(16), None, None, 0, None)


def test_ops_refuse_cpu_tensors_and_bad_layout():
    import torch
    from jerald_amd import jaxpops as J
    pos = torch.zeros(4, 3)
    with pytest.raises(L.LdlError, match="HIP device"):
        J.ppaint(pos, 1.0, (4, 4, 4), 1.0)
    with pytest.raises(Exception, match="lyidx"):
        J.ppaint(pos, 1.0, (4, 4, 4), 1.0, lyidx=5)
    with pytest.raises(NotImplementedError):
        J.initLayout(2)
    with pytest.raises(NotImplementedError, match="edges"):
        J.ppaint(pos, 1.0, (4, 4, 4), 1.0, edges=[[[0, 2]], [[0, 4]], [[0, 4]]])
    assert J.initLayout(1) == 0 and J.clean(0) == 0
    with pytest.raises(ValueError):
        J.set_vjp_rule("jax")
