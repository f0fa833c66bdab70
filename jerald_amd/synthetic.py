"""Seeded synthetic inputs for the benchmarks and studies, built on the GPU
with torch and libldl's own FP64 transforms (input setup only, not part of
any timed path).

* ``zeldovich_gpu``: lattice + Zel'dovich displacement, the recipe of the
  headline configuration: white noise n_k, delta_k = sqrt(P(k)) n_k with
  P(k) = k^-2 exp(-k^2 (2 cells)^2), psi = c2r(i k / k^2 delta_k), scaled to
  an rms of ``rms_cells`` lattice cells per axis.
* ``clumped_positions``: the precision study's variant: a seeded fraction of
  the particles moved into seeded Gaussian clumps (radii log-uniform in mesh
  cells).
"""
from __future__ import annotations

import math

import torch

from . import jaxpops as jpo


def zeldovich_gpu(n, box, rms_cells, seed, device):
    """(n^3, 3) float64 positions in box units, lattice order (z fastest)."""
    g = torch.Generator(device=device).manual_seed(seed)
    white = torch.randn((n, n, n), generator=g, device=device, dtype=torch.float64)
    wk = jpo.pfft(white)  # unnormalised r2c
    del white
    cell = box / n
    k1 = torch.fft.fftfreq(n, d=cell, device=device, dtype=torch.float64) * 2 * math.pi
    kz = k1[: n // 2 + 1]
    k2 = k1[:, None, None] ** 2 + k1[None, :, None] ** 2 + kz[None, None, :] ** 2
    k2s = torch.where(k2 > 0, k2, torch.ones_like(k2))
    pk = torch.where(k2 > 0, torch.exp(-k2 * (2 * cell) ** 2) / k2s, torch.zeros_like(k2))
    dk = torch.sqrt(pk) * wk / k2s
    del wk, pk
    psi = torch.empty((n ** 3, 3), device=device, dtype=torch.float64)
    for d, kk in enumerate((k1[:, None, None], k1[None, :, None], kz[None, None, :])):
        psi[:, d] = jpo.pifft(1j * kk * dk, (n, n, n)).reshape(-1) / float(n) ** 3  # normalised c2r
    del dk, k2, k2s
    psi *= rms_cells * cell / torch.sqrt(torch.mean(psi ** 2))
    i = torch.arange(n, device=device, dtype=torch.float64) * cell
    q = torch.stack(torch.meshgrid(i, i, i, indexing="ij"), dim=-1).reshape(-1, 3)
    return q + psi


def clumped_positions(X, box, mesh_cell, nclump=2000, frac=0.4, rmin=0.5, rmax=4.0, seed=86):
    """Move a seeded ``frac`` of the particles of X (box units) into ``nclump``
    Gaussian clumps with centres uniform in the box and radii log-uniform in
    [rmin, rmax] mesh cells; positions wrapped into [0, box)."""
    dev = X.device
    g = torch.Generator(device=dev).manual_seed(seed)
    npart = X.shape[0]
    centre = torch.rand((nclump, 3), generator=g, device=dev, dtype=torch.float64) * box
    radius = mesh_cell * torch.exp(math.log(rmin) + (math.log(rmax) - math.log(rmin))
                                   * torch.rand(nclump, generator=g, device=dev, dtype=torch.float64))
    moved = torch.rand(npart, generator=g, device=dev, dtype=torch.float64) < frac
    idx = torch.nonzero(moved).reshape(-1)
    which = torch.randint(0, nclump, (idx.numel(),), generator=g, device=dev)
    off = torch.randn((idx.numel(), 3), generator=g, device=dev, dtype=torch.float64) * radius[which, None]
    Y = X.clone()
    Y[idx] = centre[which] + off
    return torch.remainder(Y, box)
