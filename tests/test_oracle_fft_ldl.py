"""Pins the FFT conventions and the LDL model restatement of oracle/:
brute-force DFT, Parseval, adjoint identities, the reference VJP rules'
compositions, finite differences and an independent torch float64 autograd
restatement.  CPU only."""
import numpy as np
import pytest
import torch

from oracle import ldl_oracle as O

import torch_ldl as TL


def brute_r2c(x):
    n0, n1, n2 = x.shape
    out = np.zeros((n0, n1, n2 // 2 + 1), complex)
    idx = np.stack(np.meshgrid(*[np.arange(n) for n in x.shape], indexing="ij"), -1)
    for k0 in range(n0):
        for k1 in range(n1):
            for k2 in range(n2 // 2 + 1):
                ph = -2j * np.pi * (k0 * idx[..., 0] / n0 + k1 * idx[..., 1] / n1 + k2 * idx[..., 2] / n2)
                out[k0, k1, k2] = np.sum(x * np.exp(ph))
    return out


def test_pfft_brute_force_and_pifft():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((4, 6, 8))
    X = O.pfft(x)
    np.testing.assert_allclose(X, brute_r2c(x), atol=1e-10)
    np.testing.assert_allclose(O.pifft(X, x.shape), x * x.size, atol=1e-10)
    # Parseval on the half spectrum with the Hermitian weights
    c = O.hermitian_weight(8)[None, None, :]
    assert abs(np.sum(c * np.abs(X) ** 2) / x.size - np.sum(x ** 2)) < 1e-9


def test_pifft_ignores_imag_of_self_conjugate_planes():
    rng = np.random.default_rng(1)
    X = rng.standard_normal((4, 4, 5)) + 1j * rng.standard_normal((4, 4, 5))
    a = O.pifft(X, (4, 4, 8))
    # closed form: Re(sum_H c X e^{+ik.n})
    idx = np.stack(np.meshgrid(*[np.arange(n) for n in (4, 4, 8)], indexing="ij"), -1)
    c = O.hermitian_weight(8)
    b = np.zeros((4, 4, 8))
    for k0 in range(4):
        for k1 in range(4):
            for k2 in range(5):
                ph = 2j * np.pi * (k0 * idx[..., 0] / 4 + k1 * idx[..., 1] / 4 + k2 * idx[..., 2] / 8)
                b += np.real(c[k2] * X[k0, k1, k2] * np.exp(ph))
    np.testing.assert_allclose(a, b, atol=1e-10)


def test_exact_adjoints_and_reference_rules():
    rng = np.random.default_rng(2)
    shape = (4, 6, 8)
    x = rng.standard_normal(shape)
    Y = rng.standard_normal((4, 6, 5)) + 1j * rng.standard_normal((4, 6, 5))
    g = rng.standard_normal(shape)
    # <Y, pfft x> = <pfft^T Y, x>, <g, pifft Y> = <pifft^T g, Y>
    lhs = np.real(np.sum(np.conj(Y) * O.pfft(x)))
    assert abs(lhs - np.sum(O.exact_pfft_bwd(Y, shape) * x)) < 1e-9
    lhs = np.sum(g * O.pifft(Y, shape))
    assert abs(lhs - np.real(np.sum(np.conj(O.exact_pifft_bwd(g)) * Y))) < 1e-9
    # the reference's pifft rule is 1/c of the exact one, its pfft rule c times
    c = O.hermitian_weight(8)[None, None, :]
    np.testing.assert_allclose(O.ref_rule_pifft_bwd(g) * c, O.exact_pifft_bwd(g), atol=1e-10)
    np.testing.assert_allclose(O.ref_rule_pfft_bwd(Y, shape), O.exact_pfft_bwd(Y * c, shape), atol=1e-10)
    # composition pfft_bwd(pifft_bwd(.)) is exact (c and 1/c cancel) on a real kernel
    K = rng.random((4, 6, 5))
    exact = O.exact_pfft_bwd(K * O.exact_pifft_bwd(g), shape)
    ref = O.ref_rule_pfft_bwd(K * O.ref_rule_pifft_bwd(g), shape)
    np.testing.assert_allclose(ref, exact, atol=1e-9)
    # c2f doubles interior k_z planes of a real cotangent (model.py:23-29)
    v = rng.standard_normal((4, 4, 3))
    np.testing.assert_allclose(O.ref_rule_c2f_bwd(v), v * O.hermitian_weight(4)[None, None, :])


def _setup(N=8, nstep=2, seed=83, L=40.0):
    X = O.lattice_positions(N, L) + O.zeldovich_displacement(N, L, 1.5, seed)
    rng = np.random.default_rng(seed + 1)
    tgt = rng.random((N, N, N)) + 0.5
    p = np.array([0.7, 1.2, 0.3, 1.4, 0.2] * nstep + [1.1, 0.9, 0.05])
    return X, tgt, p, L


@pytest.mark.parametrize("smoothing,L1", [(None, True), ("gauss", False), (None, False)])
def test_ldl_param_grad_finite_difference(smoothing, L1):
    X, tgt, p, L = _setup()
    sm = ("gaussian", L / 8) if smoothing == "gauss" else None
    m = O.SMModelOracle(8, L, 2, tgt, L1=L1, smoothing=sm)
    loss, g = m.loss_and_grad(p, X)
    for i in range(len(p)):
        h = 1e-6 * max(abs(p[i]), 1.0)
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (m.loss(pp, X) - m.loss(pm, X)) / (2 * h)
        assert abs(g[i] - fd) <= 1e-5 * abs(fd) + 1e-9, (i, g[i], fd)


@pytest.mark.parametrize("smoothing,L1", [(None, False), ("gauss", True)])
def test_dm_param_and_position_grad_finite_difference(smoothing, L1):
    """dm kind: kernel -exp(-k/kl - kh/k) k^nu, map = density, 10 parameters."""
    X, tgt, p, L = _setup()
    p = p[:10]
    sm = ("gaussian", L / 8) if smoothing == "gauss" else None
    m = O.SMModelOracle(8, L, 2, tgt, L1=L1, smoothing=sm, kind="dm")
    loss, g, gX = m.loss_and_grad(p, X, want_pos_grad=True)
    assert g.shape == (10,)
    for i in range(len(p)):
        h = 1e-6 * max(abs(p[i]), 1.0)
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (m.loss(pp, X) - m.loss(pm, X)) / (2 * h)
        assert abs(g[i] - fd) <= 1e-5 * abs(fd) + 1e-9, (i, g[i], fd)
    for pi, d in [(5, 0), (123, 2)]:
        Xp, Xm = X.copy(), X.copy()
        Xp[pi, d] += 1e-6
        Xm[pi, d] -= 1e-6
        fd = (m.loss(p, Xp) - m.loss(p, Xm)) / 2e-6
        assert abs(gX[pi, d] - fd) <= 1e-5 * abs(fd) + 1e-10


def _hi_setup(N=8, nstep=2, seed=83, L=40.0):
    X, tgt, _, L = _setup(N, nstep, seed, L)
    rng = np.random.default_rng(seed + 2)
    star = rng.random((N, N, N)) * 1.5
    star[rng.random((N, N, N)) < 0.1] = 0.0  # empty cells: starmap^gamma has a zero base
    lay = [[7.0, 1.2, 0.3, 1.4, 0.2, 5.0, 0.8, 0.2, 1.1, 0.3], [5.0, 0.9, 0.2, 1.6, -0.1, 3.0, 1.1, 0.25, 1.3, 0.1]]
    p = np.array(sum(lay[:nstep], []) + [1.1, 0.9, 0.3, 0.4, 0.05, 2.0, 0.3, 0.2, 1.3, 0.1, 0.02, 0.05])
    return X, tgt, p, L, star


@pytest.mark.parametrize("smoothing,L1", [(None, True), ("gauss", False)])
def test_hi_param_and_position_grad_finite_difference(smoothing, L1):
    """HI kind: two sources per layer (10 parameters), depletion response (12)."""
    X, tgt, p, L, star = _hi_setup()
    sm = ("gaussian", L / 8) if smoothing == "gauss" else None
    m = O.SMModelOracle(8, L, 2, tgt, L1=L1, smoothing=sm, kind="HI", starmap=star)
    loss, g, gX = m.loss_and_grad(p, X, want_pos_grad=True)
    assert g.shape == (32,)
    for i in range(len(p)):
        h = 1e-6 * max(abs(p[i]), 1.0)
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (m.loss(pp, X) - m.loss(pm, X)) / (2 * h)
        assert abs(g[i] - fd) <= 1e-5 * abs(fd) + 1e-9, (i, g[i], fd)
    for pi, d in [(5, 0), (123, 2), (300, 1)]:
        Xp, Xm = X.copy(), X.copy()
        Xp[pi, d] += 1e-6
        Xm[pi, d] -= 1e-6
        fd = (m.loss(p, Xp) - m.loss(p, Xm)) / 2e-6
        assert abs(gX[pi, d] - fd) <= 1e-5 * abs(fd) + 1e-10


def test_hi_requires_starmap_and_zero_base_convention():
    with pytest.raises(ValueError):
        O.SMModelOracle(8, 40.0, 1, np.zeros((8, 8, 8)), kind="HI")
    np.testing.assert_array_equal(O.pow_log(np.array([0.0, 1.0]), 1.3), [0.0, 0.0])
    assert O.pow_log(np.array([2.0]), 0.5)[0] == pytest.approx(np.sqrt(2) * np.log(2))


def test_ldl_position_grad_finite_difference():
    X, tgt, p, L = _setup()
    m = O.SMModelOracle(8, L, 2, tgt, L1=False)
    _, _, gX = m.loss_and_grad(p, X, want_pos_grad=True)
    for pi, d in [(3, 0), (77, 1), (300, 2)]:
        h = 1e-6
        Xp, Xm = X.copy(), X.copy()
        Xp[pi, d] += h
        Xm[pi, d] -= h
        fd = (m.loss(p, Xp) - m.loss(p, Xm)) / (2 * h)
        assert abs(gX[pi, d] - fd) <= 1e-5 * abs(fd) + 1e-10


@pytest.mark.parametrize("kind", ["sm", "HI"])
def test_reference_rule_changes_only_alpha(kind):
    if kind == "HI":
        X, tgt, p, L, star = _hi_setup()
        alpha = [0, 5, 10, 15]  # alpha1 and alpha2 multiply outside c2f (model.py:231)
    else:
        X, tgt, p, L = _setup()
        star = None
        alpha = [0, 5]
    a = O.SMModelOracle(8, L, 2, tgt, L1=False, kind=kind, starmap=star)
    b = O.SMModelOracle(8, L, 2, tgt, L1=False, grad_rule="reference", kind=kind, starmap=star)
    _, ga = a.loss_and_grad(p, X)
    _, gb = b.loss_and_grad(p, X)
    other = [i for i in range(len(p)) if i not in alpha]
    np.testing.assert_allclose(ga[other], gb[other], rtol=1e-12)
    assert np.all(np.abs(gb[alpha]) < np.abs(ga[alpha]))


# ---------------- independent torch float64 autograd restatement ----------
@pytest.mark.parametrize("kind", ["sm", "dm", "HI"])
def test_ldl_against_torch_autograd(kind):
    N, L = 8, 40.0
    X, tgt, p, _ = _setup(N=N, L=L)
    star = None
    if kind == "dm":
        p = p[:10]
    if kind == "HI":
        X, tgt, p, _, star = _hi_setup(N=N, L=L)
    m = O.SMModelOracle(N, L, 2, tgt, L1=False, kind=kind, starmap=star)
    loss, g = m.loss_and_grad(p, X)
    pt = torch.tensor(p, requires_grad=True)
    lt = TL.model_loss(pt, X, tgt, (N, N, N), L, 2, kind=kind, star=star)
    lt.backward()
    assert abs(lt.item() - loss) < 1e-12 * max(1, abs(loss))
    np.testing.assert_allclose(pt.grad.numpy(), g, rtol=1e-9, atol=1e-14)


@pytest.mark.parametrize("nmesh,box", [((8, 12, 16), (8.0, 6.0, 20.0)), ((4, 4, 6), (1.0, 2.0, 3.0))])
def test_central_difference_is_the_derivative_symbol(nmesh, box):
    """The layer applies the derivative symbol D_d (model.py:200-201) in real
    space: the transform of the 6th-order central difference over 60 h_d is
    i D_d times the transform (forward), and the inverse transform of
    -i D_d P is minus that difference of P's inverse transform.  Checked
    against the symbol of the oracle, circularly, also where the +-3 offsets
    wrap more than once around a 4-cell axis."""
    rng = np.random.default_rng(5)
    a = rng.standard_normal(nmesh)
    D = O.fd_derivative_coeff(nmesh, box)
    for d in range(3):
        h = box[d] / nmesh[d]

        def diff(v):
            sh = lambda s: np.roll(v, -s, axis=d)  # noqa: E731  (v[x + s])
            return (45 * (sh(1) - sh(-1)) - 9 * (sh(2) - sh(-2)) + (sh(3) - sh(-3))) / (60 * h)

        np.testing.assert_allclose(np.fft.rfftn(diff(a)), 1j * D[d] * np.fft.rfftn(a), rtol=1e-12, atol=1e-12)
        P = np.fft.rfftn(rng.standard_normal(nmesh))
        np.testing.assert_allclose(np.fft.irfftn(-1j * D[d] * P, s=nmesh, axes=(0, 1, 2)),
                                   -diff(np.fft.irfftn(P, s=nmesh, axes=(0, 1, 2))),
                                   rtol=1e-12, atol=1e-12)
