"""Pins the oracle's emulation of the reference's custom VJP rules
(grad_rule="reference": jaxpops_bind.py:699-701, :758-759, model.py:23-29).

1. Against autograd applying those rules mechanically (tests/torch_ldl.py),
   on cubic meshes and on meshes whose second axis differs from the third --
   where c2f's doubling of z bins [1, int(N1/2)) is not the Hermitian weight,
   including N1/2 beyond the half spectrum (the Nyquist plane is doubled too).
2. Special cases the mathematics fixes for dL/dalpha, which the rules
   weight 1 on every half-spectrum bin where the exact adjoint weights 2 on
   interior k_z planes:
   * a density varying only along z: its spectrum lives on the k_x = k_y = 0
     line, G(0) = 0 removes k_z = 0 and D_z(Nyquist) = 0 removes k_z = N/2,
     so the reference-rule dL/dalpha is exactly 1/2 of the exact one;
   * a density varying only along x: its spectrum lives on the k_z = 0 plane,
     weight 1 in both rules, so the ratio is exactly 1.
CPU only."""
import numpy as np
import pytest
import torch

from oracle import ldl_oracle as O

import torch_ldl as TL


def _problem(nm, box, nstep, kind, seed=3):
    rng = np.random.default_rng(seed)
    np_ = int(np.prod(nm))
    X = rng.random((np_, 3)) * np.asarray(box)
    tgt = rng.random(nm) + 0.5
    star = None
    if kind == "HI":
        lay = [[7.0, 1.2, 0.3, 1.4, 0.2, 5.0, 0.8, 0.2, 1.1, 0.3], [5.0, 0.9, 0.2, 1.6, -0.1, 3.0, 1.1, 0.25, 1.3, 0.1]]
        p = np.array(sum(lay[:nstep], []) + [1.1, 0.9, 0.3, 0.4, 0.05, 2.0, 0.3, 0.2, 1.3, 0.1, 0.02, 0.05])
        star = rng.random(nm) * 1.5
    else:
        p = np.array([0.7, 1.2, 0.3, 1.4, 0.2] * nstep + ([1.1, 0.9, 0.05] if kind == "sm" else []))
    return X, tgt, p, star


@pytest.mark.parametrize("nm,box", [((8, 8, 8), (40.0, 40.0, 40.0)), ((8, 4, 12), (40.0, 24.0, 60.0)),
                                    ((6, 12, 8), (30.0, 50.0, 40.0))])
@pytest.mark.parametrize("kind", ["sm", "dm", "HI"])
@pytest.mark.parametrize("smoothing", [None, 6.0])
def test_reference_rule_matches_rules_applied_by_autograd(nm, box, kind, smoothing):
    nstep = 2
    X, tgt, p, star = _problem(nm, box, nstep, kind)
    orc = O.SMModelOracle(nm, box, nstep, tgt, L1=False, kind=kind, starmap=star, grad_rule="reference",
                          smoothing=None if smoothing is None else ("gaussian", smoothing))
    loss, g = orc.loss_and_grad(p, X)
    pt = torch.tensor(p, requires_grad=True)
    lt = TL.model_loss(pt, X, tgt, nm, box, nstep, kind=kind, star=star, rules="reference", smoothing=smoothing)
    lt.backward()
    assert abs(lt.item() - loss) < 1e-12 * abs(loss)
    np.testing.assert_allclose(g, pt.grad.numpy(), rtol=1e-9, atol=1e-13 * np.abs(g).max())
    # and the exact rule against torch's own autograd on the same mesh
    orx = O.SMModelOracle(nm, box, nstep, tgt, L1=False, kind=kind, starmap=star,
                          smoothing=None if smoothing is None else ("gaussian", smoothing))
    _, gx = orx.loss_and_grad(p, X)
    pt2 = torch.tensor(p, requires_grad=True)
    TL.model_loss(pt2, X, tgt, nm, box, nstep, kind=kind, star=star, smoothing=smoothing).backward()
    np.testing.assert_allclose(gx, pt2.grad.numpy(), rtol=1e-9, atol=1e-13 * np.abs(gx).max())


def test_c2f_weight_is_hermitian_weight_on_cubic_meshes():
    for n in (4, 6, 8, 20):
        np.testing.assert_array_equal(O.c2f_weight((n, n, n)), O.hermitian_weight(n))
    # second axis shorter than the third: only bins [1, N1/2) doubled
    np.testing.assert_array_equal(O.c2f_weight((8, 4, 12)), [1, 2, 1, 1, 1, 1, 1])
    # second axis longer: the slice runs past the half spectrum, Nyquist doubled
    np.testing.assert_array_equal(O.c2f_weight((6, 12, 8)), [1, 2, 2, 2, 2])


def one_axis_problem(N, box, axis, A=0.2, d0=0.4, m=1):
    """One particle per cell of an N^3 mesh, each on its node except along
    ``axis``, where particle i sits at i + d0 + A sin(2 pi m i / N) cells
    (0 < d - A, d + A < 1: every particle stays in its cell).  The painted
    density then varies along ``axis`` only and is exact:
        rho_j = 1 - A (sin(theta j) - sin(theta (j - 1))),  theta = 2 pi m / N."""
    cell = box / N
    q = O.lattice_positions(N, box)
    i = np.round(q[:, axis] / cell).astype(int)
    X = q.copy()
    X[:, axis] = (i + d0 + A * np.sin(2 * np.pi * m * i / N)) * cell
    return X


@pytest.mark.parametrize("axis,ratio", [(2, 0.5), (0, 1.0)])
@pytest.mark.parametrize("kind", ["sm", "dm"])
def test_reference_rule_alpha_special_cases(axis, ratio, kind):
    N, box, nstep = 8, 40.0, 2
    X = one_axis_problem(N, box, axis)
    rng = np.random.default_rng(9)
    tgt = rng.random((N, N, N)) + 0.5  # 3D target: the loss cotangent is generic
    p = np.array([0.7, 1.2, 0.3, 1.4, 0.2] * nstep + ([1.1, 0.9, 0.05] if kind == "sm" else []))
    res = {}
    for rule in ("exact", "reference"):
        orc = O.SMModelOracle(N, box, nstep, tgt, L1=False, kind=kind, grad_rule=rule)
        res[rule] = orc.loss_and_grad(p, X)[1]
    ge, gr = res["exact"], res["reference"]
    for a in (0, 5):
        assert abs(ge[a]) > 1e-8
        assert abs(gr[a] / ge[a] - ratio) < 1e-9, (a, gr[a] / ge[a])
    other = [i for i in range(len(p)) if i not in (0, 5)]
    np.testing.assert_allclose(gr[other], ge[other], rtol=1e-12)
