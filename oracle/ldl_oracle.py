"""CPU oracle for the LDL forward + gradient hot path.

TEST INFRASTRUCTURE ONLY.  This module is a plain FP64 NumPy restatement of
the reference algorithm, used by ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` as the *checker*.  The product path
(``jerald_amd``) never imports it.

Pinning (what checks each function, beyond retyping its formula):
  mesh_units ............ reference CIC golden vectors (oracle/_ref, built from
                          lib/parops_fast.h) + hand cases in tests
  paint / paint3 ........ brute-force pure-Python loops, mass conservation,
                          node-exact deposit, integer-shift equivariance,
                          reference golden vectors
  readout / readout3 .... adjoint identity <paint(m),F> = <m,readout(F)>,
                          exactness on cell-linear fields, golden vectors
  readout_grad .......... central finite differences of readout inside a cell,
                          golden vectors (lreadout3D vjp mode)
  pfft / pifft .......... brute-force O(N^6) DFT at N<=8, Parseval, numpy.fft
  sm, dm and HI layer /
  HI depletion response /
  loss / loss_and_grad .. central finite differences (FP64) for every parameter
                          and for the particle positions; an independent torch
                          float64 autograd restatement in tests (all kinds)
  reference VJP rules ... parity unpinned against JAX itself (JAX is absent);
                          the emulation follows jaxpops_bind.py:699-701,
                          :758-759 and model.py:23-29 and is checked for the
                          compositions the reference relies on (tests).

Conventions (all follow the reference):
  * positions are physical (Mpc/h); mesh units are ``(pos % L) / (L / N)``
    (jaxpops_bind.py:163-164);
  * CIC: i = floor(x), f = x - i, weights (1-f, f), corners (i, i+1) mod N, no
    half-cell shift (lib/parops_fast.h:372-410); the paint output is zeroed
    before deposit (lib/parops_fast.h:526-527).  One deliberate deviation:
    floor(x) is wrapped mod N, so x == N (possible after FP rounding of the
    modulo) deposits at cell 0 instead of being dropped;
  * r2c = unnormalised FFTW r2c / N^3 and c2r = unnormalised FFTW c2r
    (pmesh.py:112, :127; lib/parfft.h:45-46);
  * parameters per layer ``[alpha, gamma, kh, kl, nu]`` then ``[mu, w, b]``
    (model.py:268-272, :339-341).
"""
from __future__ import annotations

import numpy as np

EPS_RHO = 1e-8      # model.py:193, :344 ((rho + 1e-8) ** gamma)
RELU_LOW = 1e-12    # model.py:55 (ReLU floor)
ALPHA_SCALE = 1e-2  # model.py:198 (potk = 1e-2 * alpha * rhok * greenk)
ALPHA_SCALE_HI = 1e-3  # model.py:231 (potk = 1e-3 * (alpha1 greenk1 rhok + alpha2 greenk2 starmapk))
KNORM_ZERO = 1e-8   # model.py:102 (knorms == 0 replaced by 1e-8)


# ---------------------------------------------------------------------------
# Coordinates and CIC (lib/parops_fast.h:353-481, 615-821)
# ---------------------------------------------------------------------------
def mesh_units(pos, box, nmesh, dtype=np.float64):
    """``pos % BoxSize / (BoxSize / Nmesh)`` in ``dtype`` (jaxpops_bind.py:163-164)."""
    dtype = np.dtype(dtype)
    pos = np.asarray(pos, dtype=dtype)
    box = np.broadcast_to(np.asarray(box, dtype=dtype), (3,))
    nm = np.broadcast_to(np.asarray(nmesh), (3,))
    cell = (box / nm.astype(dtype)).astype(dtype)
    x = np.mod(pos, box).astype(dtype)
    return (x / cell).astype(dtype)


def cic_corners(x, nmesh):
    """Lower corner index (wrapped) and fractional offset per particle
    (lib/parops_fast.h:373-381 with the mod-N wrap of floor(x))."""
    nm = np.asarray(nmesh, dtype=np.int64)
    fl = np.floor(x)
    f = x - fl
    i0 = np.mod(fl.astype(np.int64), nm)
    return i0, f


def _corner_iter():
    for ox in (0, 1):
        for oy in (0, 1):
            for oz in (0, 1):
                yield ox, oy, oz


def paint(x, mass, nmesh):
    """CIC paint of particles at mesh-unit positions ``x`` (lpaint,
    lib/parops_fast.h:353-411).  ``mass`` is a scalar or an (Np,) array
    (scalar broadcast as in jaxpops_bind.py:161)."""
    nm = tuple(int(n) for n in nmesh)
    x = np.asarray(x, dtype=np.float64)
    np_ = x.shape[0]
    m = np.broadcast_to(np.asarray(mass, dtype=np.float64), (np_,))
    i0, f = cic_corners(x, nm)
    t = 1.0 - f
    out = np.zeros(nm, dtype=np.float64)
    for ox, oy, oz in _corner_iter():
        w = (f[:, 0] if ox else t[:, 0]) * (f[:, 1] if oy else t[:, 1]) * (f[:, 2] if oz else t[:, 2])
        ix = (i0[:, 0] + ox) % nm[0]
        iy = (i0[:, 1] + oy) % nm[1]
        iz = (i0[:, 2] + oz) % nm[2]
        np.add.at(out, (ix, iy, iz), w * m)
    return out


def paint3(x, mass3, nmesh):
    """Three-component paint (lpaint3D, lib/parops_fast.h:615-676): mass
    (Np,3) -> (3, N0, N1, N2)."""
    mass3 = np.asarray(mass3, dtype=np.float64)
    return np.stack([paint(x, mass3[:, d], nmesh) for d in range(3)], axis=0)


def _readout_weights(f, t, deriv_dim, cellinv):
    """Per-corner weight arrays; the ``deriv_dim`` axis uses (-1/cell, +1/cell)
    as in lreadout's vjp mode (lib/parops_fast.h:448-451, :753-766)."""
    ws = []
    for ox, oy, oz in _corner_iter():
        fac = []
        for d, o in enumerate((ox, oy, oz)):
            if d == deriv_dim:
                fac.append(np.full(f.shape[0], cellinv[d] if o else -cellinv[d]))
            else:
                fac.append(f[:, d] if o else t[:, d])
        ws.append(((ox, oy, oz), fac[0] * fac[1] * fac[2]))
    return ws


def readout(x, field, deriv_dim=-1, cellinv=(1.0, 1.0, 1.0)):
    """CIC readout (lreadout, lib/parops_fast.h:413-481).  With
    ``deriv_dim`` >= 0 returns the derivative along that axis, scaled by
    ``cellinv`` (= Nmesh / BoxSize for physical units)."""
    field = np.asarray(field, dtype=np.float64)
    nm = field.shape
    x = np.asarray(x, dtype=np.float64)
    i0, f = cic_corners(x, nm)
    t = 1.0 - f
    out = np.zeros(x.shape[0], dtype=np.float64)
    for (ox, oy, oz), w in _readout_weights(f, t, deriv_dim, cellinv):
        ix = (i0[:, 0] + ox) % nm[0]
        iy = (i0[:, 1] + oy) % nm[1]
        iz = (i0[:, 2] + oz) % nm[2]
        out += w * field[ix, iy, iz]
    return out


def readout3(x, fields3):
    """Three fields -> (Np, 3) (lreadout3D with vjp=False, lib/parops_fast.h:727-821)."""
    return np.stack([readout(x, fields3[d]) for d in range(3)], axis=-1)


def readout_grad(x, field, cellinv):
    """Gradient of the CIC interpolant of one field -> (Np, 3)
    (lreadout3D with vjp=True, lib/parops_fast.h:753-766)."""
    return np.stack([readout(x, field, d, cellinv) for d in range(3)], axis=-1)


# ---------------------------------------------------------------------------
# FFT conventions (lib/parfft.h:33-138; pmesh.py:99-127, :174-186)
# ---------------------------------------------------------------------------
def pfft(x):
    """Unnormalised r2c, half spectrum along the last axis (FFTW r2c)."""
    return np.fft.rfftn(np.asarray(x, dtype=np.float64))


def pifft(X, shape):
    """Unnormalised c2r (FFTW c2r); the imaginary parts of the k_z = 0 and
    k_z = N/2 planes are ignored after the xy transforms."""
    shape = tuple(int(s) for s in shape)
    return np.fft.irfftn(np.asarray(X, dtype=np.complex128), s=shape, axes=(0, 1, 2)) * float(np.prod(shape))


def r2c(x):
    """PMesh.r2c = pfft / N^3 (pmesh.py:112)."""
    return pfft(x) / float(np.prod(np.shape(x)))


def c2r(X, shape):
    """PMesh.c2r = pifft (pmesh.py:127)."""
    return pifft(X, shape)


def hermitian_weight(n2):
    """c(k_z): 1 on the k_z = 0 and N/2 planes, 2 on interior planes.  The
    c2r output is Re(sum_H c X e^{+ik.n}), so c2r^T = c * r2c (unnormalised)
    and r2c^T = c2r(X / c)."""
    c = np.full(n2 // 2 + 1, 2.0)
    c[0] = 1.0
    if n2 % 2 == 0:
        c[-1] = 1.0
    return c


def wavevectors(nmesh, box):
    """k_d = 2*pi*fftfreq(N_d, L_d/N_d), k_z truncated to N/2+1 entries so the
    last one is -pi/cell (pmesh.py:174-186).  Returns 3 broadcastable arrays."""
    nm = [int(n) for n in np.broadcast_to(np.asarray(nmesh), (3,))]
    bx = [float(b) for b in np.broadcast_to(np.asarray(box, dtype=np.float64), (3,))]
    kx = np.fft.fftfreq(nm[0], bx[0] / nm[0]) * 2 * np.pi
    ky = np.fft.fftfreq(nm[1], bx[1] / nm[1]) * 2 * np.pi
    kz = (np.fft.fftfreq(nm[2], bx[2] / nm[2]) * 2 * np.pi)[: nm[2] // 2 + 1]
    return kx[:, None, None], ky[None, :, None], kz[None, None, :]


def knorm(nmesh, box):
    """|k| with 0 -> 1e-8 (model.py:100-102)."""
    kx, ky, kz = wavevectors(nmesh, box)
    kn = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    return np.where(kn == 0, KNORM_ZERO, kn)


def fd_derivative_coeff(nmesh, box):
    """6th-order finite-difference gradient symbol per axis
    (45 sin kD - 9 sin 2kD + sin 3kD) / (30 D) (model.py:200-201)."""
    ks = wavevectors(nmesh, box)
    nm = np.broadcast_to(np.asarray(nmesh), (3,))
    bx = np.broadcast_to(np.asarray(box, dtype=np.float64), (3,))
    out = []
    for d in range(3):
        step = bx[d] / nm[d]
        kd = ks[d] * step
        out.append((45 * np.sin(kd) - 9 * np.sin(2 * kd) + np.sin(3 * kd)) / (30 * step))
    return out


def green_sm(kn, kh, kl, nu):
    """Stellar-mass kernel -exp(-k^2/kl^2 - kh^2/k^2) k^nu (model.py:195)."""
    return -np.exp(-kn ** 2 / kl ** 2 - kh ** 2 / kn ** 2) * kn ** nu


def green_dm(kn, kh, kl, nu):
    """Dark-matter kernel -exp(-k/kl - kh/k) k^nu (model.py:156)."""
    return -np.exp(-kn / kl - kh / kn) * kn ** nu


def pow_log(x, p):
    """x^p ln x with the value 0 where x == 0 -- the derivative of x^p in p as
    JAX forms it (log of a zero base replaced by log 1), which the reference's
    gradients of starmap^gamma2 and starmap^gamma3 inherit (model.py:228, :334)."""
    x = np.asarray(x, dtype=np.float64)
    safe = np.where(x == 0, 1.0, x)
    return x ** p * np.log(safe)


def hi_filter(kn, khi, kli, nui, xi):
    """Depletion filter exp(-k/kli - khi/k) k^nui + xi (model.py:329), and
    its derivatives in (khi, kli, nui, xi)."""
    e = np.exp(-kn / kli - khi / kn) * kn ** nui
    return e + xi, (-e / kn, e * kn / kli ** 2, e * np.log(kn), np.ones_like(kn))


def relu(x):
    """where(x > 1e-12, x, 1e-12) (model.py:54-57)."""
    return np.where(x > RELU_LOW, x, RELU_LOW)


# ---------------------------------------------------------------------------
# Loss smoothing kernels
# ---------------------------------------------------------------------------
def smoothing_kernel(nmesh, box, kind, n=0.0, sigma=None):
    """Isotropic loss weight W(k).

    kind == "reference": k^-n + 1 evaluated on the knorms array whose zeros
    were already replaced by 1e-8, so the ``knorms == 0`` guard never fires and
    W(0) = 1e8**n + 1 (model.py:102, :376-381).  n == 0 means no smoothing.
    kind == "gaussian": exp(-k^2 sigma^2 / 2) (sigma in box units)."""
    kn = knorm(nmesh, box)
    if kind == "reference":
        return kn ** (-float(n)) + 1.0
    if kind == "gaussian":
        kx, ky, kz = wavevectors(nmesh, box)
        k2 = kx ** 2 + ky ** 2 + kz ** 2
        return np.exp(-0.5 * k2 * float(sigma) ** 2)
    raise ValueError(kind)


# ---------------------------------------------------------------------------
# LDL stellar-mass model (model.py:185-212, :249-289, :291-347, :350-479)
# ---------------------------------------------------------------------------
class SMModelOracle:
    """FP64 forward and exact hand-written backward of JERALD's 'sm' kind, or
    with ``kind="dm"`` of the 'dm' kind (model.py:136-183: kernel
    -exp(-k/kl - kh/k) k^nu; the map is the density itself, model.py:346-347;
    5 parameters per layer and none for the response), or with ``kind="HI"``
    of the 'HI' kind (model.py:214-247, :313-336; needs ``starmap``):
    10 parameters per layer [alpha1, gamma1, kh1, kl1, nu1, alpha2, gamma2,
    kh2, kl2, nu2], potential 1e-3 (alpha1 G1 r2c((rho+1e-8)^gamma1) +
    alpha2 G2 r2c(starmap^gamma2)) with G1 the sm kernel and G2 = -G1's form
    at (kh2, kl2, nu2), and 12 response parameters [mu, w, b, beta1, khi, kli,
    nui, beta2, gamma3, xi, eta, zs]:
        dep = ReLU(beta1 c2r(r2c(starmap) W) + eta) + beta2 starmap^gamma3,
        W   = exp(-k/kli - khi/k) k^nui + xi,
        F   = ReLU(w ReLU(rho + zs)^mu - dep + b).

    loss options mirror JERALDModel: ``L1`` (the caller passes it; the
    reference defaults to True for sm and False for dm, model.py:115-122),
    ``n`` (reference smoothing exponent), ``index``, ``field2``, ``mask``.
    ``smoothing`` = None | "reference" | ("gaussian", sigma)."""

    def __init__(self, nmesh, box, nstep, target, L1=True, n=0.0, smoothing=None,
                 index=1.0, field2=None, mask=None, grad_rule="exact", backend="numpy", threads=0, kind="sm",
                 starmap=None):
        if kind not in ("sm", "dm", "HI"):
            raise ValueError(kind)
        self.kind = kind
        self.npot = 10 if kind == "HI" else 5          # model.py:130-133
        self.nresp = {"sm": 3, "dm": 0, "HI": 12}[kind]  # model.py:313-344
        if kind == "HI" and starmap is None:
            raise ValueError("Star map required for HI target")  # model.py:128
        self.starmap = None if starmap is None else np.asarray(starmap, dtype=np.float64)
        self.nmesh = tuple(int(v) for v in np.broadcast_to(np.asarray(nmesh), (3,)))
        self.box = np.broadcast_to(np.asarray(box, dtype=np.float64), (3,)).copy()
        self.nstep = int(nstep)
        self.M = float(np.prod(self.nmesh))
        self.target = np.asarray(target, dtype=np.float64)
        self.L1 = bool(L1)
        self.index = float(index)
        self.field2 = None if field2 is None else np.asarray(field2, dtype=np.float64)
        self.mask = None if mask is None else np.asarray(mask, dtype=np.float64)
        if smoothing is None and n > 0:
            smoothing = "reference"
        self.smoothing = smoothing
        if smoothing is None:
            self.W = None
        elif smoothing == "reference":
            self.W = smoothing_kernel(self.nmesh, self.box, "reference", n=n)
        else:
            self.W = smoothing_kernel(self.nmesh, self.box, "gaussian", sigma=smoothing[1])
        self.kn = knorm(self.nmesh, self.box)
        self.D = fd_derivative_coeff(self.nmesh, self.box)
        self.c = hermitian_weight(self.nmesh[2])[None, None, :]
        # weight of the kernel-parameter partials: c (exact adjoint), or under
        # the reference rules 1 everywhere times c2f's doubling of z bins
        # [1, int(N1/2)) (model.py:23-29: sz = v.shape[1] is the SECOND axis)
        self.cg = self.c if grad_rule == "exact" else c2f_weight(self.nmesh)[None, None, :]
        self.cellinv = np.asarray(self.nmesh, dtype=np.float64) / self.box
        if grad_rule not in ("exact", "reference"):
            raise ValueError(grad_rule)
        self.grad_rule = grad_rule
        # backend "numpy": the restatement above; "native": the same algorithm
        # with the C CIC loops (oracle/cic_oracle.c, OpenMP) and pocketfft with
        # ``threads`` workers -- used as the timed CPU baseline.
        if backend not in ("numpy", "native"):
            raise ValueError(backend)
        self.backend = backend
        self.threads = int(threads)

    # -- pieces -------------------------------------------------------------
    def _x(self, X):
        return mesh_units(X, self.box, self.nmesh)

    def _paint(self, x, m):
        if self.backend == "native":
            from . import native
            return native.paint_f64(x, m, self.nmesh, self.threads)
        return paint(x, m, self.nmesh)

    def _readout(self, x, F):
        if self.backend == "native":
            from . import native
            return native.readout_f64(x, F, -1, (1.0, 1.0, 1.0), self.threads)
        return readout(x, F)

    def _readout_grad(self, x, F):
        if self.backend == "native":
            from . import native
            return native.readout_grad_f64(x, F, self.cellinv, self.threads)
        return readout_grad(x, F, self.cellinv)

    def _rfftn(self, x):
        if self.backend == "native":
            import scipy.fft
            return scipy.fft.rfftn(x, workers=self.threads or -1)
        return np.fft.rfftn(x)

    def _irfftn(self, X):
        if self.backend == "native":
            import scipy.fft
            return scipy.fft.irfftn(X, s=self.nmesh, workers=self.threads or -1)
        return np.fft.irfftn(X, s=self.nmesh, axes=(0, 1, 2))

    def layer_forward(self, X, p, fact):
        alpha, gamma, kh, kl, nu = p[:5]
        x = self._x(X)
        rho = self._paint(x, 1.0) * fact
        u = (rho + EPS_RHO) ** gamma
        U = self._rfftn(u) / self.M
        G = (green_dm if self.kind == "dm" else green_sm)(self.kn, kh, kl, nu)
        cache = dict(x=x, rho=rho, u=u, U=U, G=G)
        if self.kind == "HI":  # model.py:221-231
            alpha2, gamma2, kh2, kl2, nu2 = p[5:10]
            S2 = self._rfftn(self.starmap ** gamma2) / self.M
            G2 = -green_sm(self.kn, kh2, kl2, nu2)
            P = ALPHA_SCALE_HI * (alpha * G * U + alpha2 * G2 * S2)
            cache.update(S2=S2, G2=G2)
        else:
            P = ALPHA_SCALE * alpha * U * G
        F = [self._irfftn(-1j * self.D[d] * P) * self.M for d in range(3)]
        s = np.stack([self._readout(x, F[d]) for d in range(3)], axis=-1)
        cache["F"] = F
        return X + s, cache

    def displace(self, params, X, keep=False):
        params = np.asarray(params, dtype=np.float64)
        fact = self.M / X.shape[0]
        caches = []
        X = np.asarray(X, dtype=np.float64)
        k = self.npot
        for i in range(self.nstep):
            X, cache = self.layer_forward(X, params[k * i: k * i + k], fact)
            caches.append(cache)
        return (X, caches) if keep else X

    def evaluate(self, params, X, keep=False):
        params = np.asarray(params, dtype=np.float64)
        fact = self.M / X.shape[0]
        Xf, caches = self.displace(params, X, keep=True)
        xf = self._x(Xf)
        rho = self._paint(xf, 1.0) * fact
        if self.kind == "dm":
            if keep:
                return rho, dict(caches=caches, xf=xf, rho=rho, F=rho, fact=fact)
            return rho
        if self.kind == "HI":
            r = self.hi_response(params[10 * self.nstep: 10 * self.nstep + 12], rho)
            r.update(caches=caches, xf=xf, rho=rho, fact=fact)
            return (r["F"], r) if keep else r["F"]
        mu, w, b = params[5 * self.nstep: 5 * self.nstep + 3]
        base = (rho + EPS_RHO) ** mu
        y = w * base + b
        F = relu(y)
        if keep:
            return F, dict(caches=caches, xf=xf, rho=rho, base=base, y=y, F=F, fact=fact)
        return F

    def hi_depletion(self, q):
        """dep = ReLU(beta1 S + eta) + beta2 starmap^gamma3 with S = c2r(r2c(
        starmap) W) (model.py:328-332); q = [beta1, khi, kli, nui, beta2,
        gamma3, xi, eta]."""
        beta1, khi, kli, nui, beta2, gamma3, xi, eta = q
        Sk = self._rfftn(self.starmap) / self.M
        W, dW = hi_filter(self.kn, khi, kli, nui, xi)
        S = self._irfftn(Sk * W) * self.M
        A = beta1 * S + eta
        s3 = self.starmap ** gamma3
        return dict(Sk=Sk, dW=dW, S=S, A=A, s3=s3, dep=relu(A) + beta2 * s3)

    def hi_response(self, q, rho):
        """F = ReLU(w ReLU(rho + zs)^mu - dep + b) (model.py:332-334);
        q = [mu, w, b, beta1, khi, kli, nui, beta2, gamma3, xi, eta, zs]."""
        mu, w, b = q[:3]
        zs = q[11]
        r = self.hi_depletion(q[3:11])
        R = relu(rho + zs)
        base = R ** mu
        y = w * base - r["dep"] + b
        r.update(R=R, base=base, y=y, F=relu(y))
        return r

    def hi_response_backward(self, q, st, g):
        """Gradient of the HI response: returns (g_rho, grad of q)."""
        mu, w, b, beta1, khi, kli, nui, beta2, gamma3, xi, eta, zs = q
        gq = np.zeros(12)
        g_y = g * (st["y"] > RELU_LOW)
        R = st["R"]
        gq[0] = np.sum(g_y * w * st["base"] * np.log(R))
        gq[1] = np.sum(g_y * st["base"])
        gq[2] = np.sum(g_y)
        g_rho = g_y * w * mu * R ** (mu - 1.0) * (st["rho"] + zs > RELU_LOW)
        gq[11] = np.sum(g_rho)
        g_dep = -g_y
        gq[7] = np.sum(g_dep * st["s3"])
        gq[8] = np.sum(g_dep * beta2 * pow_log(self.starmap, gamma3))
        g_A = g_dep * (st["A"] > RELU_LOW)
        gq[10] = np.sum(g_A)
        gq[3] = np.sum(g_A * st["S"])
        # S = M irfftn(Sk W): dL/dW = c Re(conj(rfftn(g_S)) Sk) on the half spectrum
        gW = self.cg * np.real(np.conj(self._rfftn(beta1 * g_A)) * st["Sk"])
        for j, dW in zip((4, 5, 6, 9), st["dW"]):
            gq[j] = np.sum(gW * dW)
        return g_rho, gq

    def _residual(self, F):
        Fi = F ** self.index
        if self.field2 is not None:
            Fi = Fi * self.field2
        diff = Fi - self.target
        if self.W is not None:
            diff = self._irfftn(self._rfftn(diff) * self.W)
        return diff

    def _reduce(self, diff):
        a = np.abs(diff)
        if self.mask is not None:
            a = a * self.mask
            npix = float(np.sum(self.mask))
        else:
            npix = float(diff.size)
        return (np.sum(a) if self.L1 else np.sum(a ** 2)) / npix, npix

    def loss(self, params, X):
        F = self.evaluate(params, X)
        return self._reduce(self._residual(F))[0]

    def loss_and_grad(self, params, X, want_pos_grad=False):
        params = np.asarray(params, dtype=np.float64)
        F, st = self.evaluate(params, X, keep=True)
        diff = self._residual(F)
        loss, npix = self._reduce(diff)
        grad = np.zeros_like(params)
        # d loss / d diff (abs then L1/L2, mask, 1/Npix)
        m = self.mask if self.mask is not None else 1.0
        g = (np.sign(diff) if self.L1 else 2.0 * diff) * m / npix
        if self.W is not None:
            # diff_s = C(W R(diff) / M) -> g_diff = C(W R(g)) / M
            g = self._irfftn(self._rfftn(g) * self.W)
        if self.field2 is not None:
            g = g * self.field2
        if self.index != 1.0:
            g = g * self.index * F ** (self.index - 1.0)
        if self.kind == "dm":
            g_rho = g
        elif self.kind == "HI":
            g_rho, gq = self.hi_response_backward(params[10 * self.nstep: 10 * self.nstep + 12], st, g)
            grad[10 * self.nstep: 10 * self.nstep + 12] = gq
        else:
            g_y = g * (st["y"] > RELU_LOW)
            mu, w, b = params[5 * self.nstep: 5 * self.nstep + 3]
            rhoe = st["rho"] + EPS_RHO
            grad[5 * self.nstep + 0] = np.sum(g_y * w * st["base"] * np.log(rhoe))
            grad[5 * self.nstep + 1] = np.sum(g_y * st["base"])
            grad[5 * self.nstep + 2] = np.sum(g_y)
            g_rho = g_y * w * mu * rhoe ** (mu - 1.0)
        fact = st["fact"]
        gX = fact * self._readout_grad(st["xf"], g_rho)
        k = self.npot
        for i in reversed(range(self.nstep)):
            gX = self._layer_backward(params[k * i: k * i + k], st["caches"][i], gX, fact, grad, k * i)
        return (loss, grad, gX) if want_pos_grad else (loss, grad)

    def _layer_backward(self, p, cache, v, fact, grad, off):
        alpha, gamma, kh, kl, nu = p[:5]
        x = cache["x"]
        gX = v.copy()
        gF = [self._paint(x, np.ascontiguousarray(v[:, d])) for d in range(3)]
        for d in range(3):
            gX += v[:, d:d + 1] * self._readout_grad(x, cache["F"][d])
        Q = sum(1j * self.D[d] * self._rfftn(gF[d]) for d in range(3))
        U, G = cache["U"], cache["G"]
        cw = self.c if self.grad_rule == "exact" else 1.0
        kn = self.kn
        re = np.real(np.conj(Q) * U)
        asc = ALPHA_SCALE_HI if self.kind == "HI" else ALPHA_SCALE
        grad[off + 0] = asc * np.sum(cw * G * re)
        gG = asc * alpha * self.cg * re
        if self.kind == "dm":  # d/dkh, d/dkl of -kn/kl - kh/kn
            grad[off + 2] = np.sum(gG * G * (-1.0 / kn))
            grad[off + 3] = np.sum(gG * G * (kn / kl ** 2))
        else:
            grad[off + 2] = np.sum(gG * G * (-2.0 * kh / kn ** 2))
            grad[off + 3] = np.sum(gG * G * (2.0 * kn ** 2 / kl ** 3))
        grad[off + 4] = np.sum(gG * G * np.log(kn))
        if self.kind == "HI":  # second source alpha2 G2 r2c(starmap^gamma2)
            alpha2, gamma2, kh2, kl2, nu2 = p[5:10]
            G2, S2 = cache["G2"], cache["S2"]
            re2 = np.real(np.conj(Q) * S2)
            grad[off + 5] = asc * np.sum(cw * G2 * re2)
            gG2 = asc * alpha2 * self.cg * re2
            grad[off + 7] = np.sum(gG2 * G2 * (-2.0 * kh2 / kn ** 2))
            grad[off + 8] = np.sum(gG2 * G2 * (2.0 * kn ** 2 / kl2 ** 3))
            grad[off + 9] = np.sum(gG2 * G2 * np.log(kn))
            h = self._irfftn(asc * alpha2 * G2 * Q)
            grad[off + 6] = np.sum(h * pow_log(self.starmap, gamma2))
        g_u = self._irfftn(asc * alpha * G * Q)
        rhoe = cache["rho"] + EPS_RHO
        grad[off + 1] = np.sum(g_u * cache["u"] * np.log(rhoe))
        g_rho = g_u * gamma * rhoe ** (gamma - 1.0)
        gX += fact * self._readout_grad(x, g_rho)
        return gX


# ---------------------------------------------------------------------------
# Reference VJP rules, restated in torch's gradient convention
# ---------------------------------------------------------------------------
# JAX cotangents of complex values are the complex conjugates of torch's
# gradients.  Translating the reference's custom rules:
#   pfft_bwd  (jaxpops_bind.py:699-701):  pifft(conj v).real  -> torch: c2r_unnorm(g)
#   pifft_bwd (jaxpops_bind.py:758-759):  conj(pfft(v))        -> torch: r2c_unnorm(g)
#   c2f_bwd   (model.py:23-29):           v[:, :, 1:N1/2] += conj(v[:, :, 1:N1/2])
def ref_rule_pfft_bwd(g, shape):
    return pifft(g, shape)


def ref_rule_pifft_bwd(g):
    return pfft(g)


def ref_rule_c2f_bwd(v):
    v = np.array(v, copy=True)
    sz = v.shape[1]
    v[:, :, 1:sz // 2] += np.conj(v[:, :, 1:sz // 2])
    return v


def c2f_weight(nmesh):
    """Half-spectrum weights that c2f's backward (model.py:23-29) applies to
    the cotangent of a real kernel: 2 on z bins [1, int(N1/2)) -- N1 the second
    mesh axis, the slice clipped to the N2/2+1 bins -- and 1 elsewhere (= the
    Hermitian weight c when N1 == N2)."""
    w = np.ones(int(nmesh[2]) // 2 + 1)
    w[1:int(nmesh[1]) // 2] = 2.0
    return w


def exact_pfft_bwd(g, shape):
    """r2c_unnorm^T(g) = c2r_unnorm(g / c)."""
    return pifft(np.asarray(g) / hermitian_weight(shape[2])[None, None, :], shape)


def exact_pifft_bwd(g):
    """c2r_unnorm^T(g) = c * r2c_unnorm(g)."""
    g = np.asarray(g)
    return pfft(g) * hermitian_weight(g.shape[2])[None, None, :]


# ---------------------------------------------------------------------------
# Synthetic inputs (seeded): lattice + Zel'dovich displacement, targets
# ---------------------------------------------------------------------------
def lattice_positions(nside, box, nmesh=None):
    """Uniform lattice q = i * L / nside (no half-cell offset), (Np, 3) ordered
    with the z index fastest."""
    nside = int(nside)
    box = np.broadcast_to(np.asarray(box, dtype=np.float64), (3,))
    i = np.arange(nside)
    g = np.stack(np.meshgrid(i, i, i, indexing="ij"), axis=-1).reshape(-1, 3).astype(np.float64)
    return g * (box / nside)


def zeldovich_displacement(nside, box, rms_cells, seed, ncell=None):
    """Displacement psi (Np,3) in box units: white noise n_k, delta_k =
    sqrt(P(k)) n_k with P = k^-2 exp(-k^2 (2 cells)^2), psi = c2r(i k/k^2 delta_k),
    rescaled so that rms |psi| per axis is ``rms_cells`` cells of the lattice."""
    nside = int(nside)
    box = np.broadcast_to(np.asarray(box, dtype=np.float64), (3,))
    rng = np.random.default_rng(seed)
    white = rng.standard_normal((nside, nside, nside))
    wk = np.fft.rfftn(white)
    kx, ky, kz = wavevectors((nside,) * 3, box)
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    cell = box / nside
    k2s = np.where(k2 > 0, k2, 1.0)
    pk = np.where(k2 > 0, np.exp(-k2 * (2 * cell[0]) ** 2) / k2s, 0.0)
    dk = np.sqrt(pk) * wk
    psi = np.stack([np.fft.irfftn(1j * kk / k2s * dk, s=(nside,) * 3, axes=(0, 1, 2)) for kk in (kx, ky, kz)], axis=-1)
    psi = psi.reshape(-1, 3)
    rms = np.sqrt(np.mean(psi ** 2))
    return psi * (rms_cells * cell[0] / rms)


def default_params(nstep):
    """Reference-order parameters of the 64^3 one-layer FP64 workload: alpha=0.01,
    gamma=1, kh(low-k cutoff)=0.1, kl(high-k cutoff)=1.0, nu=0 per layer; mu=1, w=1, b=0."""
    return np.array([0.01, 1.0, 0.1, 1.0, 0.0] * nstep + [1.0, 1.0, 0.0], dtype=np.float64)
