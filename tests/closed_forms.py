"""Closed forms of the LDL forward on a density with a known spectrum, to pin
the oracle's and the kernels' formulas (G(k), the 1e-2 alpha factor, the
derivative symbol D(k), the response and the losses, model.py:185-212,
:338-344, :350-403) without restating them in code.

Configuration: one particle per cell of an N^3 mesh, on its node along two
axes; along ``axis`` particle i sits at i + d0 + A sin(theta i) cells,
theta = 2 pi m / N, with A < d0 < 1 - A so every particle stays in cell i.
CIC puts weight 1 - f on cell i and f on i + 1, so the painted density is
exactly, with phi_j = theta (j - 1/2):
    rho_j = 1 - A (sin(theta j) - sin(theta (j - 1))) = 1 - a cos(phi_j),
    a = 2 A sin(theta / 2).
A field u_j = c_0 + sum_r c_r cos(r phi_j) (r m < N/2) has r2c/N^3 bins
c_r/2 exp(-i r theta/2) at k = r k_1 on the ``axis`` line, k_1 = 2 pi m / L,
and the layer's force along ``axis``,
    F_j = c2r(-i D(k) 1e-2 alpha G(k) U)_j
        = sum_r 1e-2 alpha G(r k_1) D(r k_1) c_r sin(r phi_j),
with G(k) = -exp(-k^2/kl^2 - kh^2/k^2) k^nu, D(k) = (45 sin(k h) - 9 sin(2 k h)
+ sin(3 k h)) / (30 h), h = L/N, and the k = 0 bin removed by G(0) = 0 (kh > 0).
The forces across the other axes vanish.  The particle moves by the CIC
interpolation (1 - f_i) F_i + f_i F_{i+1}.
"""
import math

import numpy as np


def lattice(N, box):
    i = np.arange(N)
    g = np.stack(np.meshgrid(i, i, i, indexing="ij"), axis=-1).reshape(-1, 3).astype(np.float64)
    return g * (box / N)


def one_axis_positions(N, box, axis, A=0.2, d0=0.4, m=1):
    cell = box / N
    X = lattice(N, box)
    i = np.round(X[:, axis] / cell).astype(np.int64)
    X[:, axis] = (i + d0 + A * np.sin(2 * math.pi * m * i / N)) * cell
    return X


def density_profile(N, A=0.2, m=1):
    """(rho_j, a, phi_j) of the painted density along the displaced axis."""
    th = 2 * math.pi * m / N
    j = np.arange(N)
    phi = th * (j - 0.5)
    a = 2 * A * math.sin(th / 2)
    return 1.0 - a * np.cos(phi), a, phi


def green(k, kh, kl, nu):
    return -math.exp(-k * k / (kl * kl) - kh * kh / (k * k)) * k ** nu


def dsym(k, h):
    return (45 * math.sin(k * h) - 9 * math.sin(2 * k * h) + math.sin(3 * k * h)) / (30 * h)


def gamma_modes(gamma, a):
    """(c_0, {r: c_r}) of u = (1 + 1e-8 - a cos phi)^gamma for gamma in {1, 2}."""
    b = 1.0 + 1e-8
    if gamma == 1:
        return b, {1: -a}
    if gamma == 2:
        return b * b + a * a / 2, {1: -2 * a * b, 2: a * a / 2}
    raise ValueError("closed form only for gamma 1 and 2")


def layer_closed_form(N, box, axis, prm, A=0.2, d0=0.4, m=1):
    """Positions (box units) after one sm layer with prm = [alpha, gamma, kh,
    kl, nu], starting from ``one_axis_positions``."""
    alpha, gamma, kh, kl, nu = prm
    X = one_axis_positions(N, box, axis, A, d0, m)
    _, a, phi = density_profile(N, A, m)
    _, cr = gamma_modes(gamma, a)
    h = box / N
    k1 = 2 * math.pi * m / box
    F = np.zeros(N)
    for r, c in cr.items():
        F += 1e-2 * alpha * green(r * k1, kh, kl, nu) * dsym(r * k1, h) * c * np.sin(r * phi)
    i = np.round(lattice(N, box)[:, axis] / h).astype(np.int64)  # each particle's cell along the axis
    f = d0 + A * np.sin(2 * math.pi * m * i / N)                   # and its CIC fraction there
    out = X.copy()
    out[:, axis] += (1 - f) * F[i % N] + f * F[(i + 1) % N]
    return out


def sm_loss_closed_form(N, mu, w, b, t, A=0.2, m=1, L1=False, smooth=None, box=None):
    """Loss of the sm response ReLU(w (rho + 1e-8)^mu + b) against a constant
    target t, the density from ``one_axis_positions`` (no layer).  mu in
    {1, 2}; the response must stay above the ReLU floor.  smooth: None,
    ("gaussian", sigma) -> W = exp(-k^2 sigma^2 / 2), or ("reference", n) ->
    W = k^-n + 1 with W(0) = 1e8^n + 1 (model.py:102, :376-381).
    Returns (loss, dL/dw, dL/db) for L2; (loss, None, None) for L1."""
    _, a, _ = density_profile(N, A, m)
    c0, cr = gamma_modes(mu, a)
    # residual y - t = d0 + sum_r w c_r cos(r phi)
    d0 = w * c0 + b - t
    amps = {r: w * c for r, c in cr.items()}
    W0, Wr = 1.0, {r: 1.0 for r in amps}
    if smooth is not None:
        k1 = 2 * math.pi * m / box
        if smooth[0] == "gaussian":
            Wr = {r: math.exp(-0.5 * (r * k1) ** 2 * smooth[1] ** 2) for r in amps}
        else:
            n = smooth[1]
            W0 = 1e8 ** n + 1.0
            Wr = {r: (r * k1) ** (-n) + 1.0 for r in amps}
    D0 = d0 * W0
    Ar = {r: amps[r] * Wr[r] for r in amps}
    if L1:
        if abs(D0) <= sum(abs(v) for v in Ar.values()):
            raise ValueError("L1 closed form needs |DC| above the modes' amplitudes")
        return abs(D0), None, None
    loss = D0 * D0 + sum(v * v for v in Ar.values()) / 2
    if smooth is not None:
        return loss, None, None
    # dL/db = 2 mean(resid); dL/dw = 2 mean(resid u^mu), u^mu = c0 + sum c_r cos
    gb = 2 * d0
    gw = 2 * (d0 * c0 + sum(amps[r] * cr[r] for r in cr) / 2)
    return loss, gw, gb
