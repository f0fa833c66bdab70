"""PMesh facade (pmesh.py:12-220) on one MI355X.

Same methods and normalisations as the reference: ``r2c = pfft / N^3``
(pmesh.py:99-112), ``c2r = pifft`` (pmesh.py:114-127), paint/readout through
the CIC operators, wavevectors ``2 pi fftfreq(N, L/N)`` with k_z truncated to
N/2+1 entries (pmesh.py:174-186).  One GPU owns the whole mesh, so
``localL = N0``, ``localS = 0`` and the edges cover the full mesh.
"""
from __future__ import annotations

import sys

import numpy as np
import torch

from . import jaxpops as jpo


class PMesh:
    def __init__(self, Nmesh, BoxSize, comm=None, device=None, dtype=torch.float32):
        self.Nmesh = np.array(np.broadcast_to(np.asarray(Nmesh, dtype=np.int64).reshape(-1), (3,)))
        self.BoxSize = np.array(np.broadcast_to(np.asarray(BoxSize, dtype=np.float64).reshape(-1), (3,)))
        self.comm = comm
        self.commrank = 0
        self.commsize = 1
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.dtype = dtype
        jpo.initLayout(self.commsize)
        lp = jpo.buildplan(self.Nmesh, comm)
        self.localL, self.localS = int(lp[0]), int(lp[1])
        self._Nmesh = tuple(int(v) for v in self.Nmesh)
        self.fftss = (self.localS, self.localS + self.localL)
        self.fftdims = (self.localL, self._Nmesh[1], self._Nmesh[2])
        self.edges = self.compute_edges()
        self.paintdims = self.fftdims

    def compute_edges(self):
        """(3, 1, 2) array of [start, stop) per axis (pmesh.py:56-66; the
        reference writes Nmesh[2] into the y stop, equal for cubic meshes)."""
        e = np.zeros((3, 1, 2), dtype=np.int64)
        e[0, 0, 1] = self.Nmesh[0]
        e[1, 0, 1] = self.Nmesh[1]
        e[2, 0, 1] = self.Nmesh[2]
        return e

    def paint(self, pos, mass, lyidx=0):
        return jpo.ppaint(pos, mass, self.Nmesh, self.BoxSize, self.edges, self.paintdims, self.comm, lyidx)

    def readout(self, pos, mesh, lyidx=0):
        return jpo.preadout(pos, mesh, self.Nmesh, self.BoxSize, self.edges, self.paintdims, self.comm, lyidx)

    def readout3D(self, pos, meshes, lyidx=0):
        return jpo.preadout3D(pos, meshes, self.Nmesh, self.BoxSize, self.edges, self.paintdims, self.comm, lyidx)

    @staticmethod
    def clean(lyidx=0):
        return jpo.clean(lyidx)

    def r2c(self, localreal):
        """Forward-normalised r2c: pfft / N^3 (pmesh.py:112)."""
        return jpo.pfft(localreal, self.fftdims, scale=1.0 / float(np.prod(self.Nmesh)))

    def c2r(self, localcplx):
        """Unnormalised c2r (pmesh.py:127)."""
        return jpo.pifft(localcplx, self.fftdims)

    def compute_wavevectors(self):
        """(N0, N1, N2/2+1, 3) wavevectors (pmesh.py:174-186)."""
        ks = []
        for d in range(3):
            n = self._Nmesh[d]
            k = np.fft.fftfreq(n, self.BoxSize[d] / n) * 2 * np.pi
            ks.append(torch.tensor(k, dtype=self.dtype, device=self.device))
        kz = ks[2][: self._Nmesh[2] // 2 + 1]
        Kx, Ky, Kz = torch.meshgrid(ks[0][self.fftss[0]:self.fftss[1]], ks[1], kz, indexing="ij")
        return torch.stack((Kx, Ky, Kz), dim=-1)

    def printr(self, strings, rank=0, flush=True):
        if self.commrank == rank:
            for s in strings if isinstance(strings, list) else [strings]:
                print(s)
            if flush:
                sys.stdout.flush()
