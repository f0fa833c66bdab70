"""Training around the LDL path: the reference's optimizers, bounds, masks,
learning-rate schedule and training loop.

* ``RMSProp`` and ``Adam`` (model.py:492-659), including the text state file
  of ``Adam.save_state`` / ``Adam.import_state``;
* ``split_masks`` (JERALD/runJERALD.py:114-127): train / validation / test
  split of the mesh by cell index;
* ``initial_params`` and ``bounds`` (runJERALD.py:145-166) and ``clip``, the
  per-step clipping of runJERALD.py:190-196;
* ``lr_schedule`` (runJERALD.py:177-180);
* ``prepare_starmap`` and ``hi_target_scale`` (runJERALD.py:82-109): the HI
  inputs' floor and normalisation;
* ``train`` (runJERALD.py:186-221): clip, ``lossv_and_grad``, keep the best
  validation parameters, Adam step.

Parameters and optimizer state live on the host as float64 NumPy vectors (a
few dozen numbers); the loss and gradient come from ``JERALDModel`` on the GPU.

    python -m jerald_amd.train --Nmesh 64 --Nstep 2 --target Mstar --Nopt 10
runs the loop on a synthetic problem (seeded Zel'dovich particles and a
target made by the model at perturbed parameters); the reference reads FastPM
particles and bigfile maps instead.
"""
from __future__ import annotations

import argparse
import time
import warnings

import numpy as np


def _lr(LR, itr):
    return LR(itr) if callable(LR) else LR


class RMSProp:
    """RMSprop (model.py:492-528)."""

    def __init__(self, LR=0.01, beta=0.999, eps=1e-8, eps_root=0.):
        self.LR, self.beta, self.eps, self.eps_root = LR, beta, eps, eps_root

    def init(self, params0):
        self.v = np.zeros(len(params0))
        self.itr = 0

    def step(self, params, gradient):
        LR = _lr(self.LR, self.itr)
        g = np.asarray(gradient, dtype=np.float64)
        self.v = (1 - self.beta) * g ** 2 + self.beta * self.v
        self.itr += 1
        vh = np.sqrt(self.v / (1. - self.beta ** self.itr) + self.eps_root)
        return np.asarray(params, dtype=np.float64) - LR * (g / (vh + self.eps))


class Adam:
    """Adam (model.py:530-659) with the reference's state file layout:
    ``mu`` / values / ``nu`` / values / ``itr`` / value / ``LR`` / value(s)
    [/ ``params`` / values], one number per line."""

    def __init__(self, LR=0.01, b1=0.9, b2=0.999, eps=1e-8, eps_root=0.):
        self.LR, self.b1, self.b2, self.eps, self.eps_root = LR, b1, b2, eps, eps_root

    def init(self, params0):
        self.mu = np.zeros(len(params0))
        self.nu = np.zeros(len(params0))
        self.itr = 0

    def step(self, params, gradient):
        LR = _lr(self.LR, self.itr)
        g = np.asarray(gradient, dtype=np.float64)
        self.mu = self.b1 * self.mu + (1. - self.b1) * g
        self.nu = self.b2 * self.nu + (1. - self.b2) * g ** 2
        self.itr += 1
        muh = self.mu / (1. - self.b1 ** self.itr)
        nuh = self.nu / (1. - self.b2 ** self.itr)
        return np.asarray(params, dtype=np.float64) - LR * (muh / (np.sqrt(nuh + self.eps_root) + self.eps))

    def get_state(self):
        return self.mu, self.nu, self.itr, _lr(self.LR, self.itr)

    def import_state(self, filename, paramsreq=False):
        """Reads a state file; with ``paramsreq`` returns (found, params)."""
        sec = read_state_sections(filename)
        self.mu = np.array(sec.get("mu", []), dtype=np.float64)
        self.nu = np.array(sec.get("nu", []), dtype=np.float64)
        self.itr = int(sec["itr"][0]) if sec.get("itr") else 0
        if paramsreq:
            if "params" in sec:
                return True, np.array(sec["params"], dtype=np.float64)
            warnings.warn(f"Parameters not found in {filename}")
            return False, 0

    def save_state(self, filename, params=None):
        sections = [("mu", self.mu), ("nu", self.nu), ("itr", [self.itr]), ("LR", _lr(self.LR, self.itr))]
        if params is not None:
            sections.append(("params", params))
        write_state_sections(filename, sections)


# Optimizer state files (the layout of model.py:589-659): each section is a
# keyword line followed by its values, one number per line.
STATE_KEYWORDS = ("mu", "nu", "itr", "LR", "params")


def write_state_sections(filename, sections):
    """Writes [(keyword, values)] in order; integers as integers, other
    numbers in %.18e (scalars and arrays alike)."""
    out = []
    for key, vals in sections:
        if key not in STATE_KEYWORDS:
            raise ValueError(f"unknown state section {key!r}")
        out.append(key)
        for v in np.atleast_1d(np.asarray(vals)).ravel():
            out.append(str(int(v)) if np.issubdtype(np.asarray(v).dtype, np.integer) else format(float(v), ".18e"))
    with open(filename, "w") as f:
        f.write("\n".join(out) + "\n")


def read_state_sections(filename):
    """{keyword: [values]} of a state file.  Lines before the first keyword
    are ignored; numbers are floats except in the ``itr`` section."""
    sec, cur = {}, None
    with open(filename) as f:
        for raw in f:
            tok = raw.strip()
            if tok in STATE_KEYWORDS:
                cur = sec.setdefault(tok, [])
            elif tok and cur is not None:
                cur.append(float(tok))
    return sec


def lr_schedule(i, LRi=0.002, LRf=0.02, tau=100, n=1):
    """Learning rate rising from LRi to LRf with time scale tau (runJERALD.py:177-180)."""
    LRt = (LRf - LRi) * (2 / (np.exp(-i / tau) + 1) - 1) + LRi
    return LRt * np.ones(n)


def split_masks(nmesh):
    """(train, validation, test) float masks of runJERALD.py:114-127: the
    validation region is z < int(0.5 N); test is the cube x, y, z < int(0.43 N)
    (taken out of validation); training is everything outside validation."""
    N = int(nmesh)
    i = np.arange(N)
    x, y, z = np.meshgrid(i, i, i, indexing="ij")
    select_validate = z < int(0.5 * N)
    select_test = (x < int(0.43 * N)) & (y < int(0.43 * N)) & (z < int(0.43 * N))
    mask_train = ~select_validate
    mask_validate = select_validate & ~select_test
    return mask_train.astype(float), mask_validate.astype(float), select_test.astype(float)


def initial_params(target, Nstep, avgfact=1.0, zshift=0.0):
    """params0 of runJERALD.py:145-153 ('Mstar' carries a fourth, unused response entry)."""
    if target == "dm":
        p = [1, 0.5, 1, 8, 0] * Nstep
    elif target == "Mstar":
        p = [1, 0.5, 1, 8, 0] * Nstep + [1, avgfact, 0, 0]
    elif target == "HI":
        p = [1, 0.5, 1, 8, 0, 1, 0.5, 1, 8, 0] * Nstep + [1, avgfact, -.1, 0.6, 1, 8, .01, 0, .1, 1, 0., zshift]
    else:
        raise ValueError(target)
    return np.array(p, dtype=np.float64)


def prepare_starmap(starmap):
    """Stellar-mass map for the HI kind as the driver prepares it
    (runJERALD.py:84-87): floored at 1e-12, then divided by its mean."""
    s = np.asarray(starmap, dtype=np.float64)
    s = np.where(s < 1e-12, 1e-12, s)
    return s / s.mean()


def hi_target_scale(redshift):
    """(divisor of the HI target map, zshift) by redshift (runJERALD.py:98-109)."""
    table = {0: (4e-13, 0.0), 2: (4e-11, -1e-3), 5: (1e-8, -0.04)}
    if redshift not in table:
        raise ValueError("redshift must be 0, 2 or 5")
    return table[redshift]


def bounds(target, Nstep):
    """Clipping bounds of runJERALD.py:156-166 (None = unbounded)."""
    layer = [(None, None), (None, None), (0., None), (0.1, None), (None, None)]
    if target == "dm":
        return layer * Nstep
    if target == "Mstar":
        return layer * Nstep + [(0.1, None), (1e-5, None), (None, None)]
    if target == "HI":
        return (layer * 2) * Nstep + [(0.1, None), (1e-5, None), (None, None), (None, None), (0., None), (0.1, None),
                                       (None, None), (None, None), (None, None), (None, None), (None, None), (None, None)]
    raise ValueError(target)


def clip(params, bnds):
    """Clip each parameter into its bound (runJERALD.py:190-196); parameters
    beyond the bounds list are left as they are."""
    p = np.array(params, dtype=np.float64)
    for j in range(min(len(p), len(bnds))):
        lo, hi = bnds[j]
        if lo is not None and p[j] < lo:
            p[j] = lo
        if hi is not None and p[j] > hi:
            p[j] = hi
    return p


def train(model, params0, bnds, Nopt=10, optimizer=None, paramsave=None, optf=None, log=print):
    """The training loop of runJERALD.py:186-221.  Returns (best params, best
    validation loss, best step, history of (train loss, validation loss))."""
    if optimizer is None:
        optimizer = Adam(lambda i: lr_schedule(i, n=len(params0)))
        optimizer.init(params0)
    params = np.array(params0, dtype=np.float64)
    best, bestparams, beststep, hist = np.inf, params.copy(), 0, []
    t0 = time.time()
    for i in range(Nopt):
        params = clip(params, bnds)
        st = time.time()
        tl, vl, g = model.lossv_and_grad(params)
        tl, vl = float(tl), float(vl)
        g = g.detach().cpu().numpy() if hasattr(g, "detach") else np.asarray(g)
        hist.append((tl, vl))
        if vl < best:
            best, bestparams, beststep = vl, params.copy(), i + 1
            if paramsave:
                np.savetxt(paramsave, bestparams)
        log(f"Step {i + 1}:\nTrain loss = {tl:.5f}, validation loss = {vl:.5f}, max gradient = {np.max(np.abs(g)):.5f}, "
            f"completed in {time.time() - st:.3f}s\nParams:\n{params}\nGrad:\n{g}")
        params = optimizer.step(params, g)
    log(f"Finished optimization in {time.time() - t0:.0f}s.\nBest validation loss: {best}, at step {beststep}\n"
        f"Best parameters:\n{bestparams}")
    if optf and hasattr(optimizer, "save_state"):
        optimizer.save_state(optf, params)
    return bestparams, best, beststep, hist


def _synthetic_positions(N, box):
    """Seeded lattice + Zel'dovich displacement (2 cells rms), lattice order."""
    rng = np.random.default_rng(83)
    wk = np.fft.rfftn(rng.standard_normal((N, N, N)))
    k1 = 2 * np.pi * np.fft.fftfreq(N, box / N)
    kz = k1[: N // 2 + 1]
    k2 = k1[:, None, None] ** 2 + k1[None, :, None] ** 2 + kz[None, None, :] ** 2
    dk = np.where(k2 > 0, np.sqrt(np.exp(-k2 * (2 * box / N) ** 2)) / np.where(k2 > 0, k2, 1) ** 1.5, 0) * wk
    psi = np.stack([np.fft.irfftn(1j * kk * dk, s=(N, N, N)).ravel()
                    for kk in (k1[:, None, None], k1[None, :, None], kz[None, None, :])], -1)
    psi *= 2.0 * box / N / np.sqrt(np.mean(psi ** 2))
    q = np.stack(np.meshgrid(*[np.arange(N) * box / N] * 3, indexing="ij"), -1).reshape(-1, 3)
    return q + psi


def main():
    """runJERALD.py's train-then-evaluate flow.  With ``--Nbodypath`` (a
    FastPM particle file, block ``Position``), ``--refpath`` (target map,
    block ``Field``) and for HI ``--Mstarpath`` it reads the reference's
    bigfile inputs; otherwise it builds a synthetic problem."""
    import torch

    from . import bigfile_io as bio
    from .model import JERALDModel
    from .pmesh import PMesh
    ap = argparse.ArgumentParser()
    ap.add_argument("--target", default="Mstar", choices=["dm", "Mstar", "HI"])
    ap.add_argument("--Nmesh", type=int, default=64)
    ap.add_argument("--BoxSize", type=float, default=None)
    ap.add_argument("--Nstep", type=int, default=2)
    ap.add_argument("--n", type=float, default=0.0)
    ap.add_argument("--Nopt", type=int, default=10)
    ap.add_argument("--dtype", default="f32", choices=["f32", "f64"])
    ap.add_argument("--Nbodypath", default=None)
    ap.add_argument("--refpath", default=None)
    ap.add_argument("--Mstarpath", default=None)
    ap.add_argument("--redshift", type=int, default=0)
    ap.add_argument("--outpath", default=None)
    a = ap.parse_args()
    N = a.Nmesh
    box = a.BoxSize or 0.8 * N
    dt = torch.float32 if a.dtype == "f32" else torch.float64
    kind = {"dm": "dm", "Mstar": "sm", "HI": "HI"}[a.target]
    pm = PMesh(N, box, device="cuda", dtype=dt)
    if a.Nbodypath:
        X, lattice = bio.load_positions(a.Nbodypath, device="cuda", dtype=dt), None
    else:
        X, lattice = torch.tensor(_synthetic_positions(N, box)), (N, N, N)
    starmap = None
    if kind == "HI":
        if a.Mstarpath:
            starmap = prepare_starmap(bio.loadfield(a.Mstarpath)[0])
        else:  # synthetic stellar-mass map: the sm model at its initial parameters
            sm = JERALDModel(X, None, pm, Nstep=a.Nstep, kind="sm", lattice=lattice)
            starmap = prepare_starmap(sm.evaluate(torch.tensor(initial_params("Mstar", a.Nstep))).cpu().numpy())
    zshift = 0.0
    if a.refpath:
        target = bio.loadfield(a.refpath)[0]
        if kind == "HI":
            div, zshift = hi_target_scale(a.redshift)
            target = target / div
        target = torch.tensor(target)
    else:  # the model itself at perturbed parameters
        npar = {"dm": 5, "sm": 5, "HI": 10}[kind] * a.Nstep + {"dm": 0, "sm": 3, "HI": 12}[kind]
        nlay = npar - {"dm": 0, "sm": 3, "HI": 12}[kind]
        p_true = initial_params(a.target, a.Nstep)[:npar]
        p_true[:nlay] *= 1.0 + 0.2 * np.cos(np.arange(nlay))
        gen = JERALDModel(X, None, pm, Nstep=a.Nstep, kind=kind, lattice=lattice, starmap=starmap)
        target = gen.evaluate(torch.tensor(p_true))
    mtr, mva, _ = split_masks(N)
    model = JERALDModel(X, target, pm, Nstep=a.Nstep, kind=kind, n=a.n, L1=(kind != "dm"),
                        masktrain=mtr, maskvalid=mva, lattice=lattice, starmap=starmap)
    p0 = initial_params(a.target, a.Nstep, avgfact=float(torch.as_tensor(target).double().mean()), zshift=zshift)
    out = a.outpath
    stem = f"{out}/{a.target}_Nmesh{N}_Nstep{a.Nstep}_n{a.n:.2f}" if out else None
    best = train(model, p0, bounds(a.target, a.Nstep), a.Nopt, paramsave=f"{stem}_bestparam.txt" if out else None,
                 optf=f"{stem}_optstate.txt" if out else None)[0]
    if out:  # the evaluated map at the best parameters (runJERALD.py:251-258)
        bio.savefield(model.evaluate(torch.tensor(best)), f"{stem}_map", pm)


if __name__ == "__main__":
    main()
