"""Pins the CIC restatements (NumPy and C) of oracle/ against the reference's
own kernels (golden vectors from lib/parops_fast.h), brute force and
closed-form properties.  CPU only."""
import itertools

import numpy as np
import pytest

from oracle import ldl_oracle as O
from oracle import native as NV


def brute_paint(x, m, nm):
    """Triple loop over particles and the 8 corners: w = prod_d (1-f or f)."""
    out = np.zeros(nm)
    for p in range(x.shape[0]):
        i = [int(np.floor(x[p, d])) for d in range(3)]
        f = [x[p, d] - i[d] for d in range(3)]
        for o in itertools.product((0, 1), repeat=3):
            w = 1.0
            for d in range(3):
                w *= f[d] if o[d] else 1.0 - f[d]
            out[(i[0] + o[0]) % nm[0], (i[1] + o[1]) % nm[1], (i[2] + o[2]) % nm[2]] += w * m[p]
    return out


@pytest.mark.parametrize("tag", ["f64", "f32"])
def test_oracle_matches_reference_golden(golden_cic, tag):
    g = golden_cic
    dt = np.float64 if tag == "f64" else np.float32
    nm = tuple(g["nmesh"])
    x = O.mesh_units(g["pos"], g["box"], nm, dt)
    np.testing.assert_array_equal(x, g[f"x_{tag}"])
    x64 = x.astype(np.float64)
    tol = dict(rtol=0, atol=1e-12) if tag == "f64" else dict(rtol=0, atol=2e-6)
    m = g["mass"].astype(dt).astype(np.float64)
    np.testing.assert_allclose(O.paint(x64, m, nm), g[f"paint_{tag}"], **tol)
    np.testing.assert_allclose(O.paint(x64, 1.0, nm), g[f"paint_unit_{tag}"], **tol)
    f = g["field"].astype(dt).astype(np.float64)
    np.testing.assert_allclose(O.readout(x64, f), g[f"readout_{tag}"], **tol)
    ci = np.array(nm) / g["box"]
    np.testing.assert_allclose(O.readout_grad(x64, f, ci), g[f"readout_grad_{tag}"], **tol)
    f3 = g["fields3"].astype(dt).astype(np.float64)
    np.testing.assert_allclose(O.readout3(x64, f3), g[f"readout3_{tag}"], **tol)
    m3 = g["mass3"].astype(dt).astype(np.float64)
    np.testing.assert_allclose(O.paint3(x64, m3, nm), g[f"paint3_{tag}"], **tol)


def test_c_oracle_matches_reference_golden(golden_cic):
    g = golden_cic
    nm = tuple(g["nmesh"])
    x = g["x_f64"]
    np.testing.assert_allclose(NV.paint_f64(x, g["mass"], nm), g["paint_f64"], atol=1e-12)
    np.testing.assert_allclose(NV.readout_f64(x, g["field"]), g["readout_f64"], atol=1e-12)
    ci = np.array(nm) / g["box"]
    np.testing.assert_allclose(NV.readout_grad_f64(x, g["field"], ci), g["readout_grad_f64"], atol=1e-12)
    xm = np.empty_like(x)
    import ctypes
    NV.cic_lib().oracle_mesh_units_f64(g["pos"].ctypes.data_as(ctypes.c_void_p), x.shape[0],
                                       g["box"].ctypes.data_as(ctypes.c_void_p),
                                       np.asarray(nm, np.int64).ctypes.data_as(ctypes.c_void_p),
                                       xm.ctypes.data_as(ctypes.c_void_p))
    np.testing.assert_array_equal(xm, x)


def test_paint_brute_force():
    rng = np.random.default_rng(3)
    nm = (5, 4, 6)
    x = rng.random((60, 3)) * np.array(nm)
    m = rng.random(60)
    np.testing.assert_allclose(O.paint(x, m, nm), brute_paint(x, m, nm), atol=1e-13)


def test_paint_mass_conservation_and_nodes():
    rng = np.random.default_rng(4)
    nm = (8, 8, 8)
    x = rng.random((1000, 3)) * 8
    m = rng.random(1000)
    assert abs(O.paint(x, m, nm).sum() - m.sum()) < 1e-10
    # a particle exactly on node (3,5,7) deposits everything there
    out = O.paint(np.array([[3.0, 5.0, 7.0]]), 2.5, nm)
    assert out[3, 5, 7] == 2.5 and out.sum() == 2.5
    # periodic wrap from the last cell
    out = O.paint(np.array([[7.5, 0.0, 0.0]]), 1.0, nm)
    assert out[7, 0, 0] == 0.5 and out[0, 0, 0] == 0.5


def test_paint_shift_equivariance():
    rng = np.random.default_rng(5)
    nm = (8, 6, 4)
    x = rng.random((200, 3)) * np.array(nm)
    a = O.paint(x, 1.0, nm)
    b = O.paint(np.mod(x + np.array([2, 1, 3]), np.array(nm)), 1.0, nm)
    np.testing.assert_allclose(np.roll(a, (2, 1, 3), axis=(0, 1, 2)), b, atol=1e-12)


def test_readout_adjoint_and_linear_exactness():
    rng = np.random.default_rng(6)
    nm = (6, 7, 8)
    x = rng.random((300, 3)) * np.array(nm)
    m = rng.standard_normal(300)
    F = rng.standard_normal(nm)
    assert abs(np.sum(O.paint(x, m, nm) * F) - np.sum(m * O.readout(x, F))) < 1e-10
    # a field linear along x (without wrap across the last cell) is read exactly
    ii = np.arange(nm[0], dtype=float)[:, None, None] * np.ones(nm)
    xs = x.copy()
    xs[:, 0] = rng.random(300) * (nm[0] - 1)
    np.testing.assert_allclose(O.readout(xs, 2.0 * ii + 1.0), 2.0 * xs[:, 0] + 1.0, atol=1e-12)
    np.testing.assert_allclose(O.readout(x, np.full(nm, 3.25)), 3.25, atol=1e-13)


def test_readout_grad_finite_difference():
    rng = np.random.default_rng(7)
    nm = (6, 6, 6)
    box = np.array([12.0, 18.0, 6.0])
    ci = np.array(nm) / box
    x = np.floor(rng.random((50, 3)) * 6) + 0.2 + 0.6 * rng.random((50, 3))  # away from cell faces
    F = rng.standard_normal(nm)
    g = O.readout_grad(x, F, ci)
    h = 1e-6
    for d in range(3):
        e = np.zeros(3)
        e[d] = h
        fd = (O.readout(x + e, F) - O.readout(x - e, F)) / (2 * h) * ci[d]
        np.testing.assert_allclose(g[:, d], fd, rtol=1e-6, atol=1e-8)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_fixed_point_deposits_keep_sign_and_sum(dtype):
    """One particle per 2x2x2 block of cells, so every cell holds exactly one
    deposit.  Small fractions (the corner (1,1,1) gets a few units of 2^-s)
    are where a remainder corner -- rint(m 2^s) minus seven rounded weight
    products -- goes negative once the FP32 complements 1 - f sum above 1;
    the split deposits must stay >= 0, sum to rint(m 2^s) per particle and
    sit within rounding of w m 2^s."""
    rng = np.random.default_rng(21)
    nm, g = (32, 32, 32), 16
    idx = np.stack(np.meshgrid(*[np.arange(g)] * 3, indexing="ij"), -1).reshape(-1, 3)
    frac = 10.0 ** rng.uniform(-4.5, -1.0, idx.shape)
    frac[rng.random(idx.shape) < 0.3] = 0.0
    pos = (2 * idx + frac).astype(dtype)
    f = (pos - np.floor(pos)).astype(np.float64)  # the fractions the paint sees (exact)
    acc, s = NV.paint_fixed(pos, 1.0, nm, box=np.array(nm, np.float64), dtype=dtype)
    M = round(2.0 ** s)
    assert acc.min() >= 0
    blocks = acc.reshape(g, 2, g, 2, g, 2).transpose(0, 2, 4, 1, 3, 5).reshape(-1, 8)
    assert np.all(blocks.sum(1) == M)
    w = np.ones((len(f), 8))
    for k in range(8):
        for d, bit in enumerate(((k >> 2) & 1, (k >> 1) & 1, k & 1)):
            w[:, k] *= f[:, d] if bit else 1.0 - f[:, d]
    tol = M * 2.0 ** -21 + 4 if dtype == np.float32 else 4
    assert np.max(np.abs(blocks - w * M)) <= tol
    if dtype == np.float32:
        # the set exercises the failure of the remainder form (FP32 weights)
        t = np.stack([(1 - f).astype(np.float32), f.astype(np.float32)], -1).astype(np.float64)
        wk = np.stack([t[:, 0, k >> 2 & 1] * t[:, 1, k >> 1 & 1] * t[:, 2, k & 1] for k in range(8)], -1)
        rem = M - np.trunc(wk[:, :7] * M).sum(1)
        assert rem.min() < 0


def test_fixed_point_emulation_close_to_f64():
    rng = np.random.default_rng(8)
    nm = (8, 8, 8)
    box = 16.0
    pos = (rng.random((2000, 3)) * 16).astype(np.float32)
    acc, s = NV.paint_fixed(pos, 1.0, nm, box=box)
    assert acc.sum() == round(2000 * 2.0 ** s)      # exact mass conservation
    ref = O.paint(O.mesh_units(pos, box, nm, np.float32).astype(np.float64), 1.0, nm)
    np.testing.assert_allclose(NV.fixed_to_float(acc, s, np.float64), ref, atol=2.0 ** -s * 8)
    # lattice + displacement form gives the same deposit as absolute positions
    lat = NV.paint_fixed(np.zeros((512, 3), np.float32), 1.0, nm, lattice=(8, 8, 8), lattice_step=(1, 1, 1))[0]
    assert np.all(lat == round(2.0 ** NV.fixed_point_shift(512, 1.0, np.float32)))
    # FP32 deposits stay below 2^30; FP64 keeps the total below 2^61
    assert NV.fixed_point_shift(10 ** 6, 1.0, np.float32) == 29
    assert NV.fixed_point_shift(10 ** 6, 1.0, np.float64) == 61 - 1 - 20
