"""The training step around the LDL path (runJERALD.py:186-221, model.py:533-577)
on the HIP kernels against the same loop driven by the FP64 oracle: clipped
Adam steps with the train/validation masks of runJERALD.py:114-127, the
learning-rate schedule of runJERALD.py:177-180 and the Mstar parameter layout
(13 + 1 unused entries)."""
import numpy as np
import pytest
import torch

from jerald_amd import train as T
from jerald_amd.model import JERALDModel
from jerald_amd.pmesh import PMesh
from oracle import ldl_oracle as O

pytestmark = pytest.mark.gpu


class _OracleModel:
    """lossv_and_grad of the oracle: (train loss, validation loss, gradient)."""

    def __init__(self, N, box, nstep, target, mtr, mva, X):
        self.tr = O.SMModelOracle(N, box, nstep, target, L1=True, mask=mtr)
        self.va = O.SMModelOracle(N, box, nstep, target, L1=True, mask=mva)
        self.X = X

    def lossv_and_grad(self, p):
        loss, g = self.tr.loss_and_grad(p, self.X)
        return loss, self.va.loss(p, self.X), g


@pytest.mark.parametrize("lattice", [False, True])
def test_adam_training_steps_match_the_oracle_loop(lattice, tmp_path):
    N, nstep, nopt = 32, 2, 5
    box = 0.8 * N
    X = T._synthetic_positions(N, box)
    p0 = T.initial_params("Mstar", nstep, avgfact=1.0)
    p_true = p0[:13].copy()
    p_true[:10] *= 1.0 + 0.2 * np.cos(np.arange(10))
    target = O.SMModelOracle(N, box, nstep, np.zeros((N, N, N))).evaluate(p_true, X)
    target = target + 0.01 * target.std() * np.random.default_rng(85).standard_normal(target.shape)
    mtr, mva, _ = T.split_masks(N)
    bnds = T.bounds("Mstar", nstep)
    runs = {}
    for name in ("hip", "oracle"):
        if name == "hip":
            pm = PMesh(N, box, device="cuda", dtype=torch.float64)
            model = JERALDModel(torch.tensor(X), target, pm, Nstep=nstep, kind="sm", L1=True, masktrain=mtr, maskvalid=mva,
                                lattice=(N, N, N) if lattice else None)
        else:
            model = _OracleModel(N, box, nstep, target, mtr, mva, X)
        opt = T.Adam(lambda i: T.lr_schedule(i, n=len(p0)))
        opt.init(p0)
        f = tmp_path / f"{name}_opt.txt"
        best, bv, step, hist = T.train(model, p0, bnds, Nopt=nopt, optimizer=opt, optf=f, log=lambda s: None)
        found, p_last = T.Adam().import_state(f, paramsreq=True)  # the parameters after the last step
        assert found
        runs[name] = dict(best=best, bv=bv, step=step, hist=np.array(hist), mu=opt.mu.copy(), nu=opt.nu.copy(), itr=opt.itr,
                          last=p_last)
    h, o = runs["hip"], runs["oracle"]
    assert h["itr"] == o["itr"] == nopt
    np.testing.assert_allclose(h["hist"], o["hist"], rtol=1e-10)
    assert h["step"] == o["step"]
    np.testing.assert_allclose(h["best"], o["best"], rtol=1e-8, atol=1e-12)
    np.testing.assert_allclose(h["last"], o["last"], rtol=1e-8, atol=1e-12)
    np.testing.assert_allclose(h["mu"], o["mu"], rtol=1e-8, atol=1e-14)
    np.testing.assert_allclose(h["nu"], o["nu"], rtol=1e-8, atol=1e-20)
    assert h["hist"][-1, 0] < h["hist"][0, 0]  # the training loss went down
