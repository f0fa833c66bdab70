"""Independent torch float64 restatement of the LDL model (model.py:185-212,
:214-247, :291-347, :350-403), used by the oracle tests as a second path.

Two gradient modes:
  * rules=None: torch's own autograd through torch.fft -- the exact adjoint;
  * rules="reference": the FFTs and c2f are custom autograd Functions whose
    backward passes are the reference's custom VJP rules, applied mechanically
    by autograd in the same places model.py applies them:
      pfft_bwd  (jaxpops_bind.py:699-701)  pifft(conj v).real
      pifft_bwd (jaxpops_bind.py:758-759)  conj(pfft(v))
      c2f_bwd   (model.py:23-29)           v[:, :, 1:int(v.shape[1]/2)] += conj(.)
    JAX cotangents of complex values are the complex conjugates of torch's
    gradients, so in torch's convention these read c2r_unnorm(g),
    r2c_unnorm(g) and g + conj(g) on the same z bins.
"""
import numpy as np
import torch

from oracle import ldl_oracle as O


def t_paint(x, mass, nm):
    """CIC paint (lib/parops_fast.h:353-411) with index_add: exact autograd."""
    fl = torch.floor(x)
    f = x - fl
    i0 = fl.long()
    out = torch.zeros(nm[0] * nm[1] * nm[2], dtype=x.dtype)
    for o in [(a, b, c) for a in (0, 1) for b in (0, 1) for c in (0, 1)]:
        w = torch.ones_like(x[:, 0]) * mass
        idx = []
        for d in range(3):
            w = w * (f[:, d] if o[d] else 1 - f[:, d])
            idx.append((i0[:, d] + o[d]) % nm[d])
        out = out.index_add(0, (idx[0] * nm[1] + idx[1]) * nm[2] + idx[2], w)
    return out.reshape(nm)


def t_readout(x, F, nm):
    """CIC readout (lib/parops_fast.h:413-481)."""
    fl = torch.floor(x)
    f = x - fl
    i0 = fl.long()
    Ff = F.reshape(-1)
    acc = torch.zeros_like(x[:, 0])
    for o in [(a, b, c) for a in (0, 1) for b in (0, 1) for c in (0, 1)]:
        w = torch.ones_like(x[:, 0])
        idx = []
        for d in range(3):
            w = w * (f[:, d] if o[d] else 1 - f[:, d])
            idx.append((i0[:, d] + o[d]) % nm[d])
        acc = acc + w * Ff[(idx[0] * nm[1] + idx[1]) * nm[2] + idx[2]]
    return acc


class _RulePfft(torch.autograd.Function):
    """Unnormalised r2c with the reference's backward rule."""

    @staticmethod
    def forward(ctx, x):
        ctx.shape = tuple(x.shape)
        return torch.fft.rfftn(x)

    @staticmethod
    def backward(ctx, g):
        return torch.fft.irfftn(g, s=ctx.shape) * float(np.prod(ctx.shape))


class _RulePifft(torch.autograd.Function):
    """Unnormalised c2r with the reference's backward rule."""

    @staticmethod
    def forward(ctx, X, shape):
        return torch.fft.irfftn(X, s=shape) * float(np.prod(shape))

    @staticmethod
    def backward(ctx, g):
        return torch.fft.rfftn(g), None


class _RuleC2F(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v):
        return v.clone()

    @staticmethod
    def backward(ctx, g):
        g = g.clone()
        sl = slice(1, int(g.shape[1] / 2))
        if g.is_complex():
            g[:, :, sl] = g[:, :, sl] + torch.conj(g[:, :, sl])
        else:  # the cotangent of a real kernel: v + conj(v) = 2 v
            g[:, :, sl] = 2.0 * g[:, :, sl]
        return g


class Ops:
    """pfft / pifft / c2f in one of the two gradient modes."""

    def __init__(self, rules=None):
        self.rules = rules

    def pfft(self, x):
        return _RulePfft.apply(x) if self.rules == "reference" else torch.fft.rfftn(x)

    def pifft(self, X, shape):
        if self.rules == "reference":
            return _RulePifft.apply(X, shape)
        return torch.fft.irfftn(X, s=shape) * float(np.prod(shape))

    def c2f(self, v):
        return _RuleC2F.apply(v) if self.rules == "reference" else v


def model_loss(pt, X, tgt, nm, box, nstep, kind="sm", star=None, rules=None, smoothing=None):
    """L2 loss (mean over all cells) of the model, as a torch scalar
    differentiable in the parameter tensor ``pt``."""
    ops = Ops(rules)
    nm = tuple(int(v) for v in nm)
    box = np.broadcast_to(np.asarray(box, dtype=np.float64), (3,))
    M = float(np.prod(nm))
    Xt = torch.tensor(X)
    kn = torch.tensor(O.knorm(nm, box))
    hs = (nm[0], nm[1], nm[2] // 2 + 1)
    D = [torch.tensor(np.broadcast_to(d, hs).copy()) for d in O.fd_derivative_coeff(nm, box)]
    cell = torch.tensor(box / np.array(nm))
    boxt = torch.tensor(box)
    fact = M / X.shape[0]
    k = 10 if kind == "HI" else 5
    lo = torch.tensor(1e-12, dtype=torch.float64)
    st = None if star is None else torch.tensor(star)
    for i in range(nstep):
        a, gam, kh, kl, nu = pt[k * i: k * i + 5]
        x = torch.remainder(Xt, boxt) / cell
        rho = t_paint(x, 1.0, nm) * fact
        rhok = ops.pfft((rho + 1e-8) ** gam) / M
        if kind == "dm":
            G = ops.c2f(-torch.exp(-kn / kl - kh / kn) * kn ** nu)
        else:
            G = ops.c2f(-torch.exp(-kn ** 2 / kl ** 2 - kh ** 2 / kn ** 2) * kn ** nu)
        if kind == "HI":
            a2, gam2, kh2, kl2, nu2 = pt[k * i + 5: k * i + 10]
            G2 = ops.c2f(torch.exp(-kn ** 2 / kl2 ** 2 - kh2 ** 2 / kn ** 2) * kn ** nu2)
            S2 = ops.pfft(st ** gam2) / M
            P = 1e-3 * (a * G * rhok + a2 * G2 * S2)
        else:
            P = 1e-2 * a * rhok * G
        s = torch.stack([t_readout(x, ops.pifft(-1j * D[d] * P, nm), nm) for d in range(3)], -1)
        Xt = Xt + s
    x = torch.remainder(Xt, boxt) / cell
    rho = t_paint(x, 1.0, nm) * fact
    off = k * nstep
    if kind == "dm":
        F = rho
    elif kind == "HI":
        mu, w, b, beta1, khi, kli, nui, beta2, gam3, xi, eta, zs = pt[off:off + 12]
        W = ops.c2f(torch.exp(-kn / kli - khi / kn) * kn ** nui + xi)
        S = ops.pifft(ops.pfft(st) / M * W, nm)
        A = beta1 * S + eta
        dep = torch.where(A > 1e-12, A, lo) + beta2 * st ** gam3
        R = torch.where(rho + zs > 1e-12, rho + zs, lo)
        y = w * R ** mu - dep + b
        F = torch.where(y > 1e-12, y, lo)
    else:
        mu, w, b = pt[off:off + 3]
        y = w * (rho + 1e-8) ** mu + b
        F = torch.where(y > 1e-12, y, lo)
    diff = F - torch.tensor(tgt)
    if smoothing is not None:
        Wl = torch.tensor(O.smoothing_kernel(nm, box, "gaussian", sigma=smoothing))
        diff = ops.pifft(ops.pfft(diff) / M * Wl, nm)
    return torch.sum(diff ** 2) / M
