"""JERALD's LDL model -- stellar-mass ('sm'), dark-matter ('dm') and neutral
hydrogen ('HI') kinds -- on one MI355X.

Mirrors ``JERALDModel`` (model.py:62-479): ``displace``, ``evaluate``,
``loss``, ``loss_grad``, ``loss_and_grad``, ``lossv``, ``lossv_and_grad``
with the reference's parameter layout (model.py:255-289, :313-344):
  sm: ``[alpha, gamma, kh, kl, nu] * Nstep + [mu, w, b]``;
  dm: ``[alpha, gamma, kh, kl, nu] * Nstep`` (the map is the density);
  HI: ``[alpha1, gamma1, kh1, kl1, nu1, alpha2, gamma2, kh2, kl2, nu2] * Nstep
  + [mu, w, b, beta1, khi, kli, nui, beta2, gamma3, xi, eta, zs]`` with the
  stellar-mass map ``starmap`` as second source and depletion input;
and loss options (``n``, ``index``, ``L1``, ``field2``, ``masktrain``,
``maskvalid``; model.py:350-468).

The whole loss is one ``torch.autograd.Function`` whose forward runs the fused
layer kernels (libldl ``ldl_{sm,dm,hi}_layer_fwd_*``), the HI depletion map
(``ldl_hi_depletion_fwd_*``) and the response/loss kernel, and whose backward
runs the hand-written adjoint kernels (``ldl_{sm,dm,hi}_layer_bwd_*``,
``ldl_hi_depletion_bwd_*``).
Per layer the forward keeps the layer's input positions and, with
``save="all"``, its density, spectrum and force meshes; with
``save="positions"`` only the positions at layer boundaries are kept and the
backward recomputes each layer's density, spectrum and forces from them (the
memory-lean mode for 1024^3) -- except the last layer's, which the backward
consumes first, so keeping it costs no extra peak memory
(``recompute_last=True`` recomputes it too).  With ``reuse_sort`` (default)
a recomputed layer also keeps the forward's particle sort (4 bytes per
particle), so its recomputation skips the sort (same bits).

Extensions over the reference, all opt-in:
  * ``lattice=(l0, l1, l2)``: X is in lattice order and positions are held as
    FP32/FP64 displacements (mesh units) from the lattice, so CIC fractions keep
    full precision independent of the box coordinate;
  * ``smoothing=("gaussian", sigma)``: Gaussian loss weight exp(-k^2 sigma^2/2);
  * ``grad_rule``: "exact" (default) or "reference" -- the reference's custom
    FFT VJP rules weight interior k_z modes by 1 instead of 2 in dL/dalpha
    (jaxpops_bind.py:758-759 with model.py:198), and c2f's backward
    (model.py:23-29) doubles the kernel parameters' z bins [1, int(N1/2)), which
    equals the exact weighting only when N1 == N2; with N1 == N2 every gradient
    but alpha's agrees with the exact rule;
  * ``paint_mode``: "fixed" (64-bit fixed point, deterministic) or "float".
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib as L


def _as_mesh(a, nm, dtype, device):
    if a is None:
        return None
    t = torch.as_tensor(a)
    return t.to(device=device, dtype=dtype).reshape(nm).contiguous()


class JERALDModel:
    def __init__(self, X, target, pm, Nstep=2, kind="sm", n=0, index=1.0, L1=None, field2=None, masktrain=None,
                 maskvalid=None, starmap=None, *, smoothing=None, grad_rule="exact", paint_mode="fixed",
                 lattice=None, save="all", recompute_last=False, dtype=None, reuse_sort=True):
        # unknown kinds use the sm model (model.py:110-112)
        self.kind = kind if kind in ("dm", "sm", "HI") else "sm"
        self._k = {"sm": "sm", "dm": "dm", "HI": "hi"}[self.kind]  # libldl entry-point prefix
        self.npot = 10 if self.kind == "HI" else 5  # parameters per layer (model.py:130-133)
        self.nresp = {"sm": 3, "dm": 0, "HI": 12}[self.kind]  # response parameters after the layers
        self.pm = pm
        self.device = pm.device
        self.dtype = dtype if dtype is not None else pm.dtype
        self.Nstep = int(Nstep)
        self.n = float(n)
        self.index = float(index)
        self.L1 = (self.kind != "dm") if L1 is None else bool(L1)  # model.py:115-122
        self.nm = tuple(int(v) for v in pm.Nmesh)
        self.box = tuple(float(v) for v in pm.BoxSize)
        self.cell = np.array(self.box) / np.array(self.nm)
        if grad_rule not in ("exact", "reference"):
            raise ValueError("grad_rule must be 'exact' or 'reference'")
        self.alpha_rule = 0 if grad_rule == "exact" else 1
        self.paint_flags = {"fixed": L.LDL_PAINT_FIXED,
                            "direct": L.LDL_PAINT_FIXED | L.LDL_PAINT_DIRECT,
                            "float": L.LDL_PAINT_FLOAT_ATOMIC}[paint_mode]
        if save not in ("all", "positions"):
            raise ValueError("save must be 'all' or 'positions'")
        self.save = save
        # save="positions": the backward recomputes every layer but the last,
        # whose state it consumes first (the same peak memory); recompute_last
        # recomputes that one too
        self.recompute_last = bool(recompute_last)
        self.reuse_sort = bool(reuse_sort)  # backward paint reuses the forward's particle sort
        # loss smoothing: reference k^-n + 1 when n > 0 (model.py:376-381)
        if smoothing is None:
            self.smooth_kind, self.smooth_par = (1, self.n) if self.n > 0 else (0, 0.0)
        elif smoothing == "reference":
            self.smooth_kind, self.smooth_par = 1, self.n
        elif isinstance(smoothing, tuple) and smoothing[0] == "gaussian":
            self.smooth_kind, self.smooth_par = 2, float(smoothing[1])
        else:
            raise ValueError("smoothing must be None, 'reference' or ('gaussian', sigma)")
        self.target = _as_mesh(target, self.nm, self.dtype, self.device)
        self.field2 = _as_mesh(field2, self.nm, self.dtype, self.device)
        self.masktrain = _as_mesh(masktrain, self.nm, self.dtype, self.device)
        self.maskvalid = _as_mesh(maskvalid, self.nm, self.dtype, self.device)
        self.starmap = _as_mesh(starmap, self.nm, self.dtype, self.device)
        self.starmapk = None
        if self.starmap is not None:  # r2c(starmap) (model.py:125-127), kept for the depletion filter
            self.starmapk = pm.r2c(self.starmap).to(torch.complex128 if self.dtype == torch.float64
                                                   else torch.complex64).contiguous()
        elif self.kind == "HI":
            raise ValueError("Star map required for HI target")  # model.py:128
        self.lattice = None if lattice is None else tuple(int(v) for v in lattice)
        self.X = None if X is None else self._encode(X)

    # ------------------------------------------------------ positions
    def _lattice_points(self):
        """Lattice points in mesh units (float64) in particle order."""
        l0, l1, l2 = self.lattice
        step = np.array(self.nm, dtype=np.float64) / np.array(self.lattice, dtype=np.float64)
        i = torch.arange(l0 * l1 * l2, device=self.device, dtype=torch.int64)
        idx = torch.stack((i // (l1 * l2), (i // l2) % l1, i % l2), dim=1).to(torch.float64)
        return idx * torch.tensor(step, device=self.device, dtype=torch.float64)

    def _encode(self, X):
        X = torch.as_tensor(X).to(self.device)
        if X.dim() != 2 or X.shape[1] != 3:
            raise ValueError("X must be (Np, 3)")
        if self.lattice is None:
            return X.to(self.dtype).contiguous()
        if int(np.prod(self.lattice)) != X.shape[0]:
            raise ValueError("lattice sizes must multiply to the number of particles")
        cell = torch.tensor(self.cell, device=self.device, dtype=torch.float64)
        d = X.to(torch.float64) / cell - self._lattice_points()
        return d.to(self.dtype).contiguous()

    def _decode(self, P):
        if self.lattice is None:
            return P
        cell = torch.tensor(self.cell, device=self.device, dtype=torch.float64)
        return ((self._lattice_points() + P.to(torch.float64)) * cell).to(self.dtype)

    def _pd(self):
        if self.lattice is None:
            return L.pos_desc(self.box)
        step = [self.nm[d] / self.lattice[d] for d in range(3)]
        return L.pos_desc(self.box, L.LDL_POS_LATTICE, self.lattice, step)

    # ------------------------------------------------------- kernels
    def _ws(self):
        np_ = 0 if self.X is None else self.X.shape[0]
        nb = L.load().ldl_sm_workspace_bytes(L.i64x3(self.nm), int(self.dtype == torch.float64), np_, self.paint_flags)
        return L.workspaces.get(self.device, f"sm{self.dtype}", nb)

    def _layer_fwd(self, P, prm, fact, keep_bins=False, need_out=True, bins=None):
        """One layer: (positions out, rho, U, F, bins, S2); S2 (HI second
        source spectrum) is None for the other kinds; positions out is None
        without ``need_out`` (the backward's recomputation).  ``bins``: the
        sort an earlier forward of the same positions wrote, read instead of
        sorting again."""
        dt, dev, nm = self.dtype, self.device, self.nm
        rho = torch.empty(nm, dtype=dt, device=dev)
        hp = L.load().ldl_sm_spectrum_pitch(L.i64x3(nm), int(dt == torch.float64))  # 64-byte rows
        U = torch.empty(nm[0], nm[1], hp, 2, dtype=dt, device=dev)
        F = torch.empty(nm + (3,), dtype=dt, device=dev)  # forces, interleaved (x, y, z) per cell
        out = torch.empty_like(P) if need_out else None
        flags = self.paint_flags
        if bins is not None:
            flags |= L.LDL_PAINT_BINS_VALID
        elif keep_bins:
            nb = L.load().ldl_sm_bins_bytes(L.i64x3(nm), P.shape[0])
            bins = torch.empty(nb, dtype=torch.uint8, device=dev)
        ws = self._ws()
        head = (L.ptr(P), self._pd(), P.shape[0], L.i64x3(nm), L.dvec(prm), float(fact), flags)
        tail = (L.ptr(F), L.ptr(bins), L.ptr(ws), ws.numel(), L.stream_ptr(dev))
        S2 = None
        if self.kind == "HI":
            S2 = torch.empty_like(U)
            L.call(f"ldl_hi_layer_fwd_{L.sfx(dt)}", *head, L.ptr(self.starmap), L.ptr(out), L.ptr(rho), L.ptr(U), L.ptr(S2),
                   *tail)
        else:
            L.call(f"ldl_{self._k}_layer_fwd_{L.sfx(dt)}", *head, L.ptr(out), L.ptr(rho), L.ptr(U), *tail)
        return out, rho, U, F, bins, S2

    def _paint_density(self, P, fact):
        dt, dev, nm = self.dtype, self.device, self.nm
        rho = torch.empty(nm, dtype=dt, device=dev)
        nb = L.load().ldl_paint_workspace_bytes(L.i64x3(nm), 1, P.shape[0], self.paint_flags)
        ws = L.workspaces.get(dev, f"paint1", nb)
        L.call(f"ldl_ppaint_{L.sfx(dt)}", L.ptr(P), self._pd(), P.shape[0], None, 1.0, L.i64x3(nm), float(fact),
               self.paint_flags, L.ptr(rho), L.ptr(ws), ws.numel(), L.stream_ptr(dev))
        return rho

    def _depletion(self, q):
        """HI depletion map: (sfield, dep) for q = [beta1, khi, kli, nui, beta2,
        gamma3, xi, eta] (model.py:326-332)."""
        dt, dev, nm = self.dtype, self.device, self.nm
        sfield = torch.empty(nm, dtype=dt, device=dev)
        dep = torch.empty(nm, dtype=dt, device=dev)
        ws = self._ws()
        L.call(f"ldl_hi_depletion_fwd_{L.sfx(dt)}", L.ptr(self.starmap), L.ptr(self.starmapk), L.i64x3(nm), L.dvec(self.box),
               L.dvec(q), L.ptr(sfield), L.ptr(dep), L.ptr(ws), ws.numel(), L.stream_ptr(dev))
        return sfield, dep

    def _loss_kernel(self, rho, q, need_grad):
        """Response + loss: out (loss at [0], response gradients, validation
        loss at [4]), g_rho and the HI extras (sfield, g_dep)."""
        dt, dev, nm = self.dtype, self.device, self.nm
        out = torch.zeros(6, dtype=torch.float64, device=dev)
        g_rho = torch.empty(nm, dtype=dt, device=dev) if need_grad else None
        ws = self._ws()
        common = (L.ptr(rho), L.ptr(self.target), L.ptr(self.masktrain), L.ptr(self.maskvalid), L.ptr(self.field2))
        geo = (L.i64x3(nm), L.dvec(self.box))
        opts = (self.index, 1 if self.L1 else 2, self.smooth_kind, self.smooth_par, L.ptr(out), L.ptr(g_rho))
        tail = (L.ptr(ws), ws.numel(), L.stream_ptr(dev))
        extra = {}
        if self.kind == "dm":
            L.call(f"ldl_dm_loss_{L.sfx(dt)}", *common, *geo, *opts, *tail)
        elif self.kind == "HI":
            sfield, dep = self._depletion(q[3:11])
            g_dep = torch.empty(nm, dtype=dt, device=dev) if need_grad else None
            L.call(f"ldl_hi_loss_{L.sfx(dt)}", *common, L.ptr(dep), *geo, L.dvec([q[0], q[1], q[2], q[11]]), *opts,
                   L.ptr(g_dep), *tail)
            extra = dict(sfield=sfield, g_dep=g_dep)
        else:
            L.call(f"ldl_sm_loss_{L.sfx(dt)}", *common, *geo, L.dvec(q), *opts, *tail)
        return out, g_rho, extra

    # ------------------------------------------------- model pieces
    @staticmethod
    def _params(params):
        p = params.detach().to("cpu", torch.float64).numpy() if torch.is_tensor(params) else np.asarray(params, np.float64)
        return p

    def _check(self, p, Nstep):
        # trailing extra entries are accepted and get a zero gradient, as in the
        # reference (runJERALD.py:150 passes 5 Nstep + 4 for 'Mstar')
        n = self.npot * Nstep + self.nresp
        if p.ndim != 1 or p.shape[0] < n:
            raise ValueError(f"expected {n} parameters, got {p.shape}")

    def _layer_prm(self, p, i):
        return p[self.npot * i: self.npot * (i + 1)]

    def displace(self, params, X=None, Nstep=None):
        """Positions after Nstep layers (model.py:249-289), box units."""
        Nstep = self.Nstep if Nstep is None else int(Nstep)
        p = self._params(params)
        P = self.X if X is None else self._encode(X)
        fact = float(np.prod(self.nm)) / P.shape[0]
        for i in range(Nstep):
            P = self._layer_fwd(P, self._layer_prm(p, i), fact)[0]
        return self._decode(P)

    def density(self, params, X=None, Nstep=None):
        """Painted density fact * paint(X_final, 1) after Nstep layers
        (model.py:308-311; the dm kind's map)."""
        Nstep = self.Nstep if Nstep is None else int(Nstep)
        p = self._params(params)
        P = self.X if X is None else self._encode(X)
        fact = float(np.prod(self.nm)) / P.shape[0]
        for i in range(Nstep):
            P = self._layer_fwd(P, self._layer_prm(p, i), fact)[0]
        return self._paint_density(P, fact)

    def evaluate(self, params, X=None, Nstep=None):
        """Model map (model.py:291-347): sm ReLU(w (rho+1e-8)^mu + b); dm the
        density; HI ReLU(w ReLU(rho + zs)^mu - dep + b)."""
        Nstep = self.Nstep if Nstep is None else int(Nstep)
        p = self._params(params)
        P = self.X if X is None else self._encode(X)
        fact = float(np.prod(self.nm)) / P.shape[0]
        for i in range(Nstep):
            P = self._layer_fwd(P, self._layer_prm(p, i), fact)[0]
        rho = self._paint_density(P, fact)
        if self.kind == "dm":
            return rho
        q = p[self.npot * Nstep: self.npot * Nstep + self.nresp]
        F = torch.empty_like(rho)
        if self.kind == "HI":
            _, dep = self._depletion(q[3:11])
            L.call(f"ldl_hi_response_{L.sfx(self.dtype)}", L.ptr(rho), L.ptr(dep), L.i64x3(self.nm),
                   L.dvec([q[0], q[1], q[2], q[11]]), L.ptr(F), L.stream_ptr(self.device))
        else:
            L.call(f"ldl_sm_response_{L.sfx(self.dtype)}", L.ptr(rho), L.i64x3(self.nm), L.dvec(q[:3]), L.ptr(F),
                   L.stream_ptr(self.device))
        return F

    def _forward(self, p, need_grad):
        self._check(p, self.Nstep)
        P = self.X
        fact = float(np.prod(self.nm)) / P.shape[0]
        saved = []
        for i in range(self.Nstep):
            out, rho, U, F, bins, S2 = self._layer_fwd(P, self._layer_prm(p, i), fact,
                                                       keep_bins=need_grad and self.reuse_sort)
            if need_grad:
                keep = self.save == "all" or (i == self.Nstep - 1 and not self.recompute_last)
                saved.append((P, rho, U, F, bins, S2) if keep else (P, bins))
            P = out
        rho_f = self._paint_density(P, fact)
        q = p[self.npot * self.Nstep: self.npot * self.Nstep + self.nresp]
        out, g_rho, extra = self._loss_kernel(rho_f, q, need_grad)
        return out, dict(saved=saved, P=P, g_rho=g_rho, fact=fact, out5=out, q=q, **extra)

    def _backward(self, p, st):
        dt, dev, nm = self.dtype, self.device, self.nm
        P = st["P"]
        fact = st["fact"]
        g = torch.empty_like(P)
        gmax = torch.zeros(3, dtype=torch.float64, device=dev)  # max |g| per component, handed from call to call
        L.call(f"ldl_paint_vjp_pos_{L.sfx(dt)}", L.ptr(P), self._pd(), P.shape[0], L.ptr(st["g_rho"]), L.i64x3(nm),
               float(fact), L.ptr(g), L.ptr(gmax), L.stream_ptr(dev))
        k = self.npot
        grads = torch.zeros(k * self.Nstep, dtype=torch.float64, device=dev)
        ws = self._ws()
        for i in reversed(range(self.Nstep)):
            prm = self._layer_prm(p, i)
            sv = st["saved"][i]
            if len(sv) == 6:
                Pi, rho, U, F, bins, S2 = sv
            else:
                Pi, bins = sv
                _, rho, U, F, bins, S2 = self._layer_fwd(Pi, prm, fact, keep_bins=self.reuse_sort, need_out=False, bins=bins)
            gp = C_ptr_offset(grads, k * i)
            head = (L.ptr(Pi), self._pd(), Pi.shape[0], L.i64x3(nm), L.dvec(prm), float(fact), self.alpha_rule,
                    self.paint_flags)
            tail = (L.ptr(F), L.ptr(bins), L.ptr(g), L.ptr(g), gp, L.ptr(gmax), L.ptr(ws), ws.numel(), L.stream_ptr(dev))
            if self.kind == "HI":
                L.call(f"ldl_hi_layer_bwd_{L.sfx(dt)}", *head, L.ptr(self.starmap), L.ptr(rho), L.ptr(U), L.ptr(S2), *tail)
            else:
                L.call(f"ldl_{self._k}_layer_bwd_{L.sfx(dt)}", *head, L.ptr(rho), L.ptr(U), *tail)
            st["saved"][i] = None
        out = st["out5"]
        if self.kind == "HI":
            g8 = torch.empty(8, dtype=torch.float64, device=dev)
            q = st["q"]
            L.call(f"ldl_hi_depletion_bwd_{L.sfx(dt)}", L.ptr(self.starmap), L.ptr(self.starmapk), L.ptr(st["sfield"]),
                   L.ptr(st["g_dep"]), L.i64x3(nm), L.dvec(self.box), L.dvec(q[3:11]), self.alpha_rule, L.ptr(g8), L.ptr(ws), ws.numel(),
                   L.stream_ptr(dev))
            resp = torch.cat([out[1:4], g8, out[5:6]])
        else:
            resp = out[1:1 + self.nresp]
        extra = torch.zeros(len(p) - k * self.Nstep - self.nresp, dtype=torch.float64, device=dev)
        return torch.cat([grads, resp, extra]), g

    # ------------------------------------------------------- losses
    def loss(self, params):
        """Training loss (model.py:350-403) as a differentiable 0-d tensor."""
        return _LDLLoss.apply(torch.as_tensor(params, dtype=torch.float64), self)[0]

    def lossv(self, params):
        """(loss, validation loss) (model.py:415-468)."""
        p = self._params(params)
        out5, _ = self._forward(p, False)
        return out5[0], out5[4]

    def loss_grad(self, params):
        return self.loss_and_grad(params)[1]

    def loss_and_grad(self, params):
        p = self._params(params)
        out5, st = self._forward(p, True)
        grad, _ = self._backward(p, st)
        return out5[0], grad

    def lossv_and_grad(self, params):
        """(loss, validation loss, gradient) as the reference returns them
        (model.py:476-479)."""
        p = self._params(params)
        out5, st = self._forward(p, True)
        grad, _ = self._backward(p, st)
        return out5[0], out5[4], grad

    def loss_and_grad_positions(self, params):
        """Loss and dL/dX of the initial positions (box units)."""
        p = self._params(params)
        out5, st = self._forward(p, True)
        grad, g = self._backward(p, st)
        return out5[0], grad, g


def C_ptr_offset(t, n_elems):
    """Device pointer ``n_elems`` elements past the start of ``t``."""
    import ctypes
    return ctypes.c_void_p(t.data_ptr() + n_elems * t.element_size())


class _LDLLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, model):
        p = model._params(params)
        need = bool(ctx.needs_input_grad[0]) if hasattr(ctx, "needs_input_grad") else True
        out5, st = model._forward(p, need)
        ctx.model, ctx.p, ctx.st = model, p, st
        ctx.param_device, ctx.param_dtype = params.device, params.dtype
        return out5[0].clone(), out5[4].clone()

    @staticmethod
    def backward(ctx, gloss, glossv):
        grad, _ = ctx.model._backward(ctx.p, ctx.st)
        ctx.st = None
        return (grad * gloss).to(ctx.param_device, ctx.param_dtype), None
