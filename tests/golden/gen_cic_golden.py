"""Writes tests/golden/cic_ref_golden.npz: CIC golden vectors produced by the
reference's own lib/parops_fast.h kernels (compiled by `make -C oracle ref`
into oracle/_ref/libjerald_ref.so, run at MPI comm size 1).

Inputs are seeded: a non-cubic 8x6x10 mesh with box (20, 15, 30), 300
particles drawn in [-L, 2L) (exercising the periodic wrap), plus hand-placed
edge cases (exact mesh nodes, the box origin, coordinates just below L).
Positions are converted to mesh units with the reference's formula
(jaxpops_bind.py:163-164) in each precision before the call.

Run from the repository root:  python tests/golden/gen_cic_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import ldl_oracle as O  # noqa: E402
from oracle import native as NV  # noqa: E402


def main():
    rng = np.random.default_rng(2024)
    nm = (8, 6, 10)
    box = np.array([20.0, 15.0, 30.0])
    pos = rng.random((300, 3)) * box * 3 - box
    cell = box / np.array(nm)
    edge = np.array([
        [0.0, 0.0, 0.0],
        cell * np.array([3, 2, 7]),                 # exact node
        box - 1e-9,                                  # just below L (f64)
        box * (1 - 1e-7),                            # just below L (f32 ulp region)
        -cell * 0.5,                                 # negative -> wraps
        box + cell * np.array([0.25, 0.5, 0.75]),    # beyond L -> wraps
    ])
    pos = np.concatenate([pos, edge])
    mass = rng.random(pos.shape[0]) + 0.5
    mass3 = rng.standard_normal((pos.shape[0], 3))
    field = rng.standard_normal(nm)
    fields3 = rng.standard_normal((3,) + nm)
    out = dict(nmesh=np.array(nm), box=box, pos=pos, mass=mass, mass3=mass3, field=field, fields3=fields3)
    for dt, tag in ((np.float64, "f64"), (np.float32, "f32")):
        x = O.mesh_units(pos, box, nm, dt)
        out[f"x_{tag}"] = x
        out[f"paint_{tag}"] = NV.ref_paint(x, mass.astype(dt), nm, dt)
        out[f"paint_unit_{tag}"] = NV.ref_paint(x, np.ones(len(x), dt), nm, dt)
        out[f"readout_{tag}"] = NV.ref_readout(x, field.astype(dt), box, -1, dt)
        out[f"readout_grad_{tag}"] = NV.ref_readout3(x, field.astype(dt), box, True, dt)
        out[f"readout3_{tag}"] = NV.ref_readout3(x, fields3.astype(dt), box, False, dt)
        out[f"paint3_{tag}"] = NV.ref_paint3(x, mass3.astype(dt), nm, dt)
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "cic_ref_golden.npz"), **out)
    print("wrote", sorted(out))


if __name__ == "__main__":
    main()
