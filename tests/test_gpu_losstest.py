"""The reference's own loss test configuration (losstest.py: 20^3 mesh and
particles, 200 Mpc/h box, random positions with a cosine displacement, 2 sm
layers, random target, all-ones train mask and random validation mask) through
jerald_amd.losstest, against the FP64 oracle."""
import numpy as np
import pytest
import torch

from jerald_amd import losstest
from oracle import ldl_oracle as O

pytestmark = pytest.mark.gpu


def test_reference_losstest_configuration_matches_oracle():
    pos, param, target, mask, maskv, nstep = losstest.inputs(20)
    loss, lossv, grad, _, _ = losstest.run(20, torch.float64, verbose=False)
    orc = O.SMModelOracle(20, 200.0, nstep, target, L1=True, mask=mask)
    orv = O.SMModelOracle(20, 200.0, nstep, target, L1=True, mask=maskv)
    l0, g0 = orc.loss_and_grad(param, pos)
    assert abs(loss - l0) <= 1e-10 * abs(l0)
    assert abs(lossv - orv.loss(param, pos)) <= 1e-10 * abs(lossv)
    np.testing.assert_allclose(grad, g0, rtol=1e-8, atol=1e-12 * np.abs(g0).max())


def test_reference_losstest_configuration_fp32():
    """The FP32 path on the same inputs stays within 1e-5 of FP64 for the loss
    and every gradient."""
    l64, lv64, g64, _, _ = losstest.run(20, torch.float64, verbose=False)
    l32, lv32, g32, _, _ = losstest.run(20, torch.float32, verbose=False)
    assert abs(l32 - l64) <= 1e-5 * abs(l64)
    assert abs(lv32 - lv64) <= 1e-5 * abs(lv64)
    assert np.all(np.abs(g32 - g64) <= 1e-5 * np.abs(g64)), np.abs(g32 - g64) / np.abs(g64)
