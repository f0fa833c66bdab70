"""The precision study and the large single-GPU run of the LDL path.

    python -m jerald_amd.studies precision   # 512^3 particles, 1024^3 mesh, clumps
    python -m jerald_amd.studies large       # 1024^3 particles and mesh, 3 layers

Each prints one JSON line.

precision: 512^3 particles (box 409.6 Mpc/h) on a 1024^3 mesh; 40% of the
particles in 2000 seeded Gaussian clumps (radii log-uniform in [0.5, 4] mesh
cells), the rest on the Zel'dovich lattice; 2 sm layers, L2 loss with 1-cell
Gaussian smoothing, target = the model at +-10% perturbed parameters plus 1%
noise.  Runs the FP32 model with the 64-bit fixed-point paint and with
floating-point atomics, ``reruns`` times each, and the FP64 fixed-point run
of the same kernels as reference, then the FP32 model in the benchmark's
form (FP32 displacements from the 512^3 lattice, fixed-point paint).  Reports
bit-exactness across reruns (loss, gradients and the density of every rerun)
and under a particle permutation, the spread of the float-atomic reruns, and
the max/rms relative error of the density (of the input positions; for the
lattice form, after the layers), the loss and every parameter gradient
against FP64.

large: 1024^3 particles (FP32 displacements from the lattice) on a 1024^3
mesh, 3 sm layers, positions checkpointed at layer boundaries and each
layer recomputed in the backward (``save="positions"``): iterations/s and
peak HBM; the gradient error of the same configuration at 256^3 against its
FP64 run.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np
import torch

from . import _lib as L
from .model import JERALDModel
from .pmesh import PMesh
from . import jaxpops as J
from .synthetic import clumped_positions, zeldovich_gpu


def _params(nstep):
    return np.array([0.01, 1.0, 0.1, 1.0, 0.0] * nstep + [1.0, 1.0, 0.0])


def _target(X, pm, nstep, cell, p, lattice, seed=84):
    gen = JERALDModel(X, None, pm, Nstep=nstep, kind="sm", L1=False, smoothing=("gaussian", cell), lattice=lattice,
                      save="positions")
    rng = np.random.default_rng(seed)
    pt = p * (1 + 0.1 * (2 * rng.random(p.shape) - 1))
    t = gen.evaluate(torch.tensor(pt)).double()
    g = torch.Generator(device=t.device).manual_seed(seed + 1)
    t += 0.01 * t.std() * torch.randn(t.shape, generator=g, device=t.device, dtype=torch.float64)
    return t


def _rel(a, b):
    """(max |a-b| / max |b|, rms(a-b) / rms(b)) over all elements."""
    a = a.double().reshape(-1)
    b = b.double().reshape(-1)
    d = a - b
    return (float(d.abs().max() / b.abs().max()), float(d.norm() / b.norm()))


def precision(n=512, nmesh=1024, reruns=10, log=print):
    dev = torch.device("cuda:0")
    box = 0.8 * n
    cell = box / nmesh
    nstep = 2
    p = _params(nstep)
    X = clumped_positions(zeldovich_gpu(n, box, 2.0, 83, dev), box, cell)
    pm32 = PMesh(nmesh, box, device=dev, dtype=torch.float32)
    pm64 = PMesh(nmesh, box, device=dev, dtype=torch.float64)
    target = _target(X, pm64, nstep, cell, p, None)
    kw = dict(Nstep=nstep, kind="sm", L1=False, smoothing=("gaussian", cell), save="positions")
    out = {"study": "precision", "particles": n ** 3, "mesh": nmesh, "box": box, "clumps": 2000, "clump_fraction": 0.4,
           "reruns": reruns}
    # FP64 reference: the same kernels in FP64 with the fixed-point paint
    m64 = JERALDModel(X, target, pm64, **kw)
    l64, g64 = m64.loss_and_grad(p)
    rho64 = pm64.paint(X.to(torch.float64), 1.0)
    rho64_final = m64.density(torch.tensor(p))  # after the layers: what the loss sees
    del m64
    L.workspaces.clear()
    torch.cuda.empty_cache()
    res = {}
    for mode in ("fixed", "float"):
        log(f"[precision] FP32 {mode} paint, {reruns} runs")
        m = JERALDModel(X, target, pm32, paint_mode=mode, **kw)
        runs = []
        for _ in range(reruns):
            l, g = m.loss_and_grad(p)
            runs.append((l.item(), g.cpu().numpy()))
        del m
        losses = np.array([r[0] for r in runs])
        grads = np.stack([r[1] for r in runs])
        L.workspaces.clear()
        torch.cuda.empty_cache()
        # every rerun's density against the first (one extra mesh held at a time)
        J.set_paint_mode(mode)
        try:
            X32 = X.to(torch.float32)
            rho0 = pm32.paint(X32, 1.0)
            dens_same = True
            for _ in range(reruns - 1):
                dens_same &= bool(torch.equal(rho0, pm32.paint(X32, 1.0)))
            del X32
        finally:
            J.set_paint_mode("fixed")
        rhos = [rho0]
        g64n = g64.cpu().numpy()
        r = {
            "loss_bitwise_identical": bool(np.all(losses == losses[0])),
            "grad_bitwise_identical": bool(np.all(grads == grads[0])),
            "density_bitwise_identical": dens_same,
            "density_reruns": reruns,
            "loss_rerun_spread": float((losses.max() - losses.min()) / abs(l64.item())),
            "grad_rerun_spread_max": float(np.max((grads.max(0) - grads.min(0)) / np.abs(g64n))),
            "loss_rel_err": float(abs(losses[0] - l64.item()) / abs(l64.item())),
            "grad_rel_err": [float(x) for x in np.abs(grads[0] - g64n) / np.abs(g64n)],
            "density_rel_err_max_rms": _rel(rhos[0], rho64),
        }
        if mode == "fixed":  # order independence: the same particles permuted
            perm = torch.randperm(X.shape[0], device=dev, generator=torch.Generator(device=dev).manual_seed(7))
            rp = pm32.paint(X.to(torch.float32)[perm], 1.0)
            r["density_identical_under_permutation"] = bool(torch.equal(rp, rhos[0]))
        res[mode] = r
        del rhos
        torch.cuda.empty_cache()
    out["fp32_fixed_point"] = res["fixed"]
    out["fp32_float_atomic"] = res["float"]
    # the benchmark's form: FP32 displacements from the 512^3 lattice (step 2
    # cells), fixed-point paint
    log(f"[precision] FP32 lattice displacements, fixed-point paint, {reruns} runs")
    m = JERALDModel(X, target, pm32, lattice=(n, n, n), **kw)
    runs = []
    for _ in range(reruns):
        l, g = m.loss_and_grad(p)
        runs.append((l.item(), g.cpu().numpy()))
    losses = np.array([r[0] for r in runs])
    grads = np.stack([r[1] for r in runs])
    g64n = g64.cpu().numpy()
    rho_final = m.density(torch.tensor(p))
    out["fp32_lattice_fixed_point"] = {
        "loss_bitwise_identical": bool(np.all(losses == losses[0])),
        "grad_bitwise_identical": bool(np.all(grads == grads[0])),
        "loss_rel_err": float(abs(losses[0] - l64.item()) / abs(l64.item())),
        "grad_rel_err": [float(x) for x in np.abs(grads[0] - g64n) / np.abs(g64n)],
        "final_density_rel_err_max_rms": _rel(rho_final, rho64_final),
    }
    del m, rho_final
    L.workspaces.clear()
    torch.cuda.empty_cache()
    out["loss_fp64"] = l64.item()
    return out


def large(n=1024, nstep=3, steps=2, sub=256, log=print):
    dev = torch.device("cuda:0")
    box = 0.8 * n
    cell = box / n
    p = _params(nstep)
    out = {"study": "large", "particles": n ** 3, "mesh": n, "layers": nstep, "save": "positions"}
    log(f"[large] inputs n={n}")
    X = zeldovich_gpu(n, box, 2.0, 83, dev)
    pm = PMesh(n, box, device=dev, dtype=torch.float32)
    target = _target(X, pm, nstep, cell, p, (n, n, n)).float()
    m = JERALDModel(X, target, pm, Nstep=nstep, kind="sm", L1=False, smoothing=("gaussian", cell), lattice=(n, n, n),
                    save="positions")
    del X
    L.workspaces.clear()  # the input generator's FP64 transform workspace
    torch.cuda.empty_cache()
    m.loss_and_grad(p)
    torch.cuda.synchronize()
    torch.cuda.reset_peak_memory_stats(dev)
    t0 = time.perf_counter()
    for _ in range(steps):
        loss, grad = m.loss_and_grad(p)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    out.update({"value": round(1.0 / dt, 4), "unit": "it/s", "ms_per_step": round(1e3 * dt, 2),
                "peak_hbm_gb": round(torch.cuda.max_memory_allocated(dev) / 1e9, 2), "loss": loss.item()})
    del m, target
    L.workspaces.clear()
    torch.cuda.empty_cache()
    # gradient error at the sub-run size against FP64
    log(f"[large] {sub}^3 sub-run, FP32 against FP64")
    bs = 0.8 * sub
    Xs = zeldovich_gpu(sub, bs, 2.0, 83, dev)
    ts = _target(Xs, PMesh(sub, bs, device=dev, dtype=torch.float64), nstep, bs / sub, p, (sub, sub, sub))
    gs = {}
    for dt_ in (torch.float32, torch.float64):
        ms = JERALDModel(Xs, ts, PMesh(sub, bs, device=dev, dtype=dt_), Nstep=nstep, kind="sm", L1=False,
                         smoothing=("gaussian", bs / sub), lattice=(sub, sub, sub))
        l_, g_ = ms.loss_and_grad(p)
        gs[dt_] = (l_.item(), g_.cpu().numpy())
    (l32, g32), (l64, g64) = gs[torch.float32], gs[torch.float64]
    out["subrun"] = {"particles": sub ** 3, "loss_rel_err": abs(l32 - l64) / abs(l64),
                     "grad_rel_err": [float(x) for x in np.abs(g32 - g64) / np.abs(g64)]}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("which", choices=["precision", "large"])
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--mesh", type=int, default=1024)
    ap.add_argument("--reruns", type=int, default=10)
    ap.add_argument("--steps", type=int, default=2)
    a = ap.parse_args()
    log = lambda s: print(s, file=sys.stderr, flush=True)  # noqa: E731
    if a.which == "precision":
        r = precision(n=a.n or 512, nmesh=a.mesh, reruns=a.reruns, log=log)
    else:
        r = large(n=a.n or 1024, steps=a.steps, log=log)
    print(json.dumps(r))


if __name__ == "__main__":
    main()
