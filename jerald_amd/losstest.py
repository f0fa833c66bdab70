"""The reference's loss test harness (losstest.py:1-70) on the GPU path.

Same construction, in the same order and from the same NumPy seed: a
``ngrid``^3 mesh in a 200 Mpc/h box, ngrid^3 uniform random particles moved by
10 cos(|x|/19) along every axis, 2 sm layers with parameters
[.1, .5, 1., 8., 0.] * 2 + [1., 1., 0.], a uniform random target, an all-ones
train mask and a random 0/1 validation mask; then lossv once to warm up,
timed lossv, and timed lossv_and_grad with the gradient printed.  FP64, as
the reference runs with jax_enable_x64.

    python -m jerald_amd.losstest [--ngrid 20] [--dtype f64|f32]
"""
import argparse
import time

import numpy as np
import torch

from .model import JERALDModel
from .pmesh import PMesh


def inputs(ngrid=20, box=200.0):
    """Positions, parameters, target and masks of losstest.py:21-50."""
    np.random.seed(83)
    pos = np.random.rand(ngrid ** 3, 3) * box
    vel = 10 * np.cos(np.linalg.norm(pos, axis=1) / 19).reshape((len(pos), 1))
    pos = pos + vel
    nstep = 2
    param = np.array([.1, .5, 1., 8., 0.] * nstep + [1., 1., 0.])
    target = np.random.rand(ngrid, ngrid, ngrid)
    mask = np.ones_like(target)
    maskv = np.random.randint(0, 2, target.shape).astype(np.float64)
    return pos, param, target, mask, maskv, nstep


def run(ngrid=20, dtype=torch.float64, device="cuda", verbose=True):
    """Returns (loss, lossv, grad, ms_lossv, ms_lossv_and_grad)."""
    box = 200.0
    pos, param, target, mask, maskv, nstep = inputs(ngrid, box)
    pm = PMesh(ngrid, box, device=device, dtype=dtype)
    pm.printr("Mesh initialized")
    model = JERALDModel(torch.tensor(pos), target, pm, Nstep=nstep, kind="sm", masktrain=mask, maskvalid=maskv)
    p = torch.tensor(param)
    model.lossv(p)  # warm-up (the reference's JIT compilation call)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss, lossv = model.lossv(p)
    torch.cuda.synchronize()
    t_loss = (time.perf_counter() - t0) * 1e3
    if verbose:
        pm.printr([f"Loss alone computed in {t_loss:.3f}ms", f"Loss: {loss.item()}, validation loss: {lossv.item()}"])
    model.lossv_and_grad(p)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss, lossv, grad = model.lossv_and_grad(p)
    torch.cuda.synchronize()
    t_grad = (time.perf_counter() - t0) * 1e3
    if verbose:
        pm.printr([f"Loss with gradient computed in {t_grad:.3f}ms", "Gradient:", f"{grad.cpu().numpy()}"])
        pm.printr(f"Peak device memory: {torch.cuda.max_memory_allocated() / 1024:.1f} KiB")
    return loss.item(), lossv.item(), grad.cpu().numpy(), t_loss, t_grad


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ngrid", type=int, default=20)
    ap.add_argument("--dtype", default="f64", choices=["f64", "f32"])
    a = ap.parse_args()
    run(a.ngrid, torch.float64 if a.dtype == "f64" else torch.float32)


if __name__ == "__main__":
    main()
