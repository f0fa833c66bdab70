"""Pins the oracle's LDL forward formulas with closed forms (tests/
closed_forms.py): one layer's displacement of a single-mode density for
gamma 1 and 2 (G(k), 1e-2 alpha, D(k), the CIC interpolation), and the sm
response and L1/L2 losses, plain and smoothed (Gaussian, and the reference's
k^-n + 1 with W(0) = 1e8^n + 1).  A swapped kh/kl, a sign error in G or D,
a wrong power of the density or a dropped 1e-2 fails these.  CPU only."""
import math

import numpy as np
import pytest

from oracle import ldl_oracle as O

import closed_forms as CF

PRM = [40.0, 1.0, 0.3, 1.5, 0.4]  # alpha, gamma, kh (low-k cutoff), kl (high-k cutoff), nu


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("gamma,m", [(1.0, 1), (1.0, 3), (2.0, 1), (2.0, 2)])
def test_layer_displacement_closed_form(axis, gamma, m):
    N, box = 16, 24.0
    prm = list(PRM)
    prm[1] = gamma
    X = CF.one_axis_positions(N, box, axis, m=m)
    want = CF.layer_closed_form(N, box, axis, prm, m=m)
    orc = O.SMModelOracle(N, box, 1, np.ones((N, N, N)), L1=False)
    got = orc.displace(np.array(prm + [1.0, 1.0, 0.0]), X)
    moved = want[:, axis] - X[:, axis]
    assert np.abs(moved).max() > 1e-4  # the layer moves the particles
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.abs(moved).max() + 1e-14)


def test_closed_form_sees_swapped_cutoffs_and_signs():
    """The closed form separates the alternatives a misreading would give."""
    N, box, axis = 16, 24.0, 0
    base = CF.layer_closed_form(N, box, axis, PRM)
    swapped = CF.layer_closed_form(N, box, axis, [PRM[0], PRM[1], PRM[3], PRM[2], PRM[4]])
    flipped = CF.layer_closed_form(N, box, axis, [-PRM[0]] + PRM[1:])
    X = CF.one_axis_positions(N, box, axis)
    d = np.abs(base - X).max()
    assert np.abs(swapped - base).max() > 0.1 * d
    assert np.abs(flipped - base).max() > 1.5 * d


@pytest.mark.parametrize("mu", [1.0, 2.0])
@pytest.mark.parametrize("axis", [0, 2])
def test_sm_response_l2_loss_and_gradients(mu, axis):
    N, box = 16, 24.0
    w, b, t = 1.3, 0.2, 0.7
    X = CF.one_axis_positions(N, box, axis)
    orc = O.SMModelOracle(N, box, 0, np.full((N, N, N), t), L1=False)
    loss, g = orc.loss_and_grad(np.array([mu, w, b]), X)
    want, gw, gb = CF.sm_loss_closed_form(N, mu, w, b, t)
    assert abs(loss - want) <= 1e-13 * want
    assert abs(g[1] - gw) <= 1e-12 * abs(gw)
    assert abs(g[2] - gb) <= 1e-12 * abs(gb)
    # the response map itself
    rho, a, phi = CF.density_profile(N)
    prof = w * (rho + 1e-8) ** mu + b
    F = orc.evaluate(np.array([mu, w, b]), X)
    shape = [1, 1, 1]
    shape[axis] = N
    np.testing.assert_allclose(F, np.broadcast_to(prof.reshape(shape), (N, N, N)), rtol=1e-14)


@pytest.mark.parametrize("smooth", [("gaussian", 2.5), ("reference", 1.0)])
def test_smoothed_losses(smooth):
    N, box = 16, 24.0
    mu, w, b, t = 1.0, 1.3, 0.2, 0.7
    X = CF.one_axis_positions(N, box, 1)
    tgt = np.full((N, N, N), t)
    if smooth[0] == "gaussian":
        orc = O.SMModelOracle(N, box, 0, tgt, L1=False, smoothing=smooth)
        want = CF.sm_loss_closed_form(N, mu, w, b, t, smooth=smooth, box=box)[0]
    else:  # the reference's kind: L1 with k^-n + 1
        orc = O.SMModelOracle(N, box, 0, tgt, L1=True, n=smooth[1])
        want = CF.sm_loss_closed_form(N, mu, w, b, t, smooth=smooth, box=box, L1=True)[0]
    loss = orc.loss(np.array([mu, w, b]), X)
    assert abs(loss - want) <= 1e-12 * want


def test_l1_loss_closed_form():
    N, box = 16, 24.0
    X = CF.one_axis_positions(N, box, 0)
    orc = O.SMModelOracle(N, box, 0, np.full((N, N, N), 0.1), L1=True)
    want = CF.sm_loss_closed_form(N, 1.0, 1.3, 0.2, 0.1, L1=True)[0]
    assert abs(orc.loss(np.array([1.0, 1.3, 0.2]), X) - want) <= 1e-14


def test_density_profile_is_the_paint():
    N, box = 16, 24.0
    for axis in range(3):
        X = CF.one_axis_positions(N, box, axis, m=3)
        rho = O.paint(O.mesh_units(X, box, (N, N, N)), 1.0, (N, N, N))
        prof = CF.density_profile(N, m=3)[0]
        shape = [1, 1, 1]
        shape[axis] = N
        np.testing.assert_allclose(rho, np.broadcast_to(prof.reshape(shape), (N, N, N)), atol=1e-14)
    assert math.isclose(prof.mean(), 1.0, rel_tol=1e-14)
