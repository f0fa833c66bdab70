"""GPU parity of the 3D FFTs and of the fused LDL model against the oracle.

Tolerances: FP64 kernels vs the FP64 oracle to ~1e-10 relative; FP32 runs
vs the FP64 oracle to the 1e-5 relative the LDL path targets for the loss and
parameter gradients."""
import numpy as np
import pytest
import torch

from jerald_amd import jaxpops as J
from jerald_amd.model import JERALDModel
from jerald_amd.pmesh import PMesh
from oracle import ldl_oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda"

SHAPES = [(8, 8, 8), (4, 16, 8), (32, 16, 64), (64, 64, 64), (128, 32, 16), (16, 8, 256), (8, 4, 1024), (512, 4, 8)]
# axes with factors 3 and 5 take the mixed-radix kernels (and (8, 8, 4) its 2-point z pass)
MIXED = [(20, 20, 20), (12, 10, 6), (6, 24, 30), (40, 12, 18), (30, 45, 60), (8, 8, 4), (1000, 4, 10), (4, 6, 1000), (96, 80, 50)]


@pytest.mark.parametrize("shape", SHAPES + MIXED)
@pytest.mark.parametrize("dt", [torch.float32, torch.float64])
def test_pfft_pifft_match_numpy(shape, dt):
    rng = np.random.default_rng(sum(shape))
    x = rng.standard_normal(shape)
    X = np.fft.rfftn(x)
    got = J.pfft(torch.tensor(x, device=DEV, dtype=dt)).cpu().numpy()
    scale = np.abs(X).max()
    rt = 3e-6 if dt == torch.float32 else 1e-13
    assert np.abs(got - X).max() <= rt * scale * np.log2(x.size)
    # c2r of a non-Hermitian spectrum (random imaginary parts on the k_z = 0, N/2 planes)
    Y = rng.standard_normal(X.shape) + 1j * rng.standard_normal(X.shape)
    want = O.pifft(Y, shape)
    cdt = torch.complex64 if dt == torch.float32 else torch.complex128
    Yt = torch.tensor(Y, device=DEV, dtype=cdt)
    got = J.pifft(Yt, shape).cpu().numpy()
    assert np.abs(got - want).max() <= rt * np.abs(want).max() * np.log2(x.size)
    assert torch.equal(Yt, torch.tensor(Y, device=DEV, dtype=cdt))  # input untouched


def test_fft_round_trip_large():
    """Size-independent property at 256^3: c2r(r2c(x)) = N^3 x, Parseval."""
    shape = (256, 256, 256)
    x = torch.randn(shape, device=DEV, generator=torch.Generator(DEV).manual_seed(0))
    X = J.pfft(x)
    c = J.hermitian_weight(256, DEV, torch.float64)
    e_k = ((X.abs().double() ** 2) * c).sum() / x.numel()
    e_x = (x.double() ** 2).sum()
    assert abs(e_k.item() / e_x.item() - 1) < 1e-5
    y = J.pifft(X, shape) / x.numel()
    assert (y - x).abs().max().item() < 1e-5


@pytest.mark.parametrize("rule", ["exact", "reference"])
def test_fft_vjp_rules(rule):
    shape = (8, 16, 8)
    rng = np.random.default_rng(7)
    x = torch.tensor(rng.standard_normal(shape), device=DEV, dtype=torch.float64, requires_grad=True)
    G = rng.standard_normal((8, 16, 5)) + 1j * rng.standard_normal((8, 16, 5))
    J.set_vjp_rule(rule)
    try:
        J.pfft(x).backward(torch.tensor(G, device=DEV))
        want = O.exact_pfft_bwd(G, shape) if rule == "exact" else O.ref_rule_pfft_bwd(G, shape)
        np.testing.assert_allclose(x.grad.cpu().numpy(), want, rtol=1e-10, atol=1e-10)
        Y = torch.tensor(G, device=DEV, requires_grad=True)
        g = rng.standard_normal(shape)
        J.pifft(Y, shape).backward(torch.tensor(g, device=DEV))
        want = O.exact_pifft_bwd(g) if rule == "exact" else O.ref_rule_pifft_bwd(g)
        np.testing.assert_allclose(Y.grad.cpu().numpy(), want, rtol=1e-10, atol=1e-10)
        if rule == "exact":  # agrees with torch's own FFT autograd
            x2 = x.detach().clone().requires_grad_(True)
            torch.fft.rfftn(x2).backward(torch.tensor(G, device=DEV))
            np.testing.assert_allclose(x.grad.cpu().numpy(), x2.grad.cpu().numpy(), rtol=1e-10, atol=1e-10)
    finally:
        J.set_vjp_rule("reference")


def _problem(N=16, nstep=2, L=64.0, seed=83, smoothing=None, L1=True, mask=False, rms=1.5):
    X = O.lattice_positions(N, L) + O.zeldovich_displacement(N, L, rms, seed)
    rng = np.random.default_rng(seed + 1)
    tgt = rng.random((N, N, N)) + 0.5
    mk = (rng.random((N, N, N)) > 0.3).astype(float) if mask else None
    p = np.array([0.8, 1.2, 0.3, 1.5, 0.2] * nstep + [1.1, 0.9, 0.05])
    orc = O.SMModelOracle(N, L, nstep, tgt, L1=L1, smoothing=smoothing, mask=mk)
    return X, tgt, mk, p, orc


def _model(X, tgt, mk, N, L, nstep, dt, **kw):
    pm = PMesh(N, L, device=DEV, dtype=dt)
    return JERALDModel(torch.tensor(X), tgt, pm, Nstep=nstep, kind="sm", masktrain=mk, **kw)


@pytest.mark.parametrize("smoothing,L1,mask", [(None, True, True), (None, False, False),
                                               (("gaussian", 4.0), False, False), (("gaussian", 4.0), True, True)])
@pytest.mark.parametrize("lattice", [None, (16, 16, 16)])
def test_ldl_fp64_matches_oracle(smoothing, L1, mask, lattice):
    N, L, nstep = 16, 64.0, 2
    X, tgt, mk, p, orc = _problem(N, nstep, L, smoothing=smoothing, L1=L1, mask=mask)
    loss0, g0 = orc.loss_and_grad(p, X)
    m = _model(X, tgt, mk, N, L, nstep, torch.float64, L1=L1, smoothing=smoothing, lattice=lattice)
    loss, grad = m.loss_and_grad(torch.tensor(p))
    assert abs(loss.item() - loss0) <= 1e-10 * abs(loss0)
    np.testing.assert_allclose(grad.cpu().numpy(), g0, rtol=1e-8, atol=1e-12 * np.abs(g0).max())


@pytest.mark.parametrize("lat", [(4, 8, 256), (8, 2, 512), (3, 5, 256)])
def test_ldl_blocked_sweep_lattice(lat):
    """Lattices whose rows hold a multiple of 256 particles take the blocked
    (slab, row-block) gather order: same loss and gradients as the oracle."""
    N, L, nstep = 16, 64.0, 2
    rng = np.random.default_rng(7)
    idx = np.stack(np.meshgrid(*[np.arange(v) for v in lat], indexing="ij"), axis=-1).reshape(-1, 3)
    X = idx * (L / np.array(lat)) + rng.uniform(-1.5, 1.5, idx.shape) * (L / N)
    tgt = rng.random((N, N, N)) + 0.5
    p = np.array([0.8, 1.2, 0.3, 1.5, 0.2] * nstep + [1.1, 0.9, 0.05])
    orc = O.SMModelOracle(N, L, nstep, tgt, L1=False, smoothing=("gaussian", 4.0))
    loss0, g0 = orc.loss_and_grad(p, X)
    m = _model(X, tgt, None, N, L, nstep, torch.float64, L1=False, smoothing=("gaussian", 4.0), lattice=lat)
    loss, grad = m.loss_and_grad(torch.tensor(p))
    assert abs(loss.item() - loss0) <= 1e-10 * abs(loss0)
    np.testing.assert_allclose(grad.cpu().numpy(), g0, rtol=1e-8, atol=1e-12 * np.abs(g0).max())


@pytest.mark.parametrize("N,lattice", [(20, None), (12, (12, 12, 12)), (24, None)])
@pytest.mark.parametrize("smoothing,L1", [(None, True), (("gaussian", 4.0), False), ("reference", True)])
def test_ldl_mixed_radix_mesh_matches_oracle(N, lattice, smoothing, L1):
    """Meshes whose axes are not powers of two run the unfused x passes
    around the elementwise filter kernels: same loss and gradients."""
    L, nstep = 4.0 * N, 2
    X, tgt, mk, p, _ = _problem(N, nstep, L, smoothing=None, L1=L1, mask=True)
    ref = smoothing == "reference"  # k^-n + 1 with n = 1 (the model's n argument)
    orc = O.SMModelOracle(N, L, nstep, tgt, L1=L1, n=1.0 if ref else 0.0, smoothing=None if ref else smoothing, mask=mk)
    loss0, g0 = orc.loss_and_grad(p, X)
    kw = dict(n=1.0) if ref else dict(smoothing=smoothing)
    m = _model(X, tgt, mk, N, L, nstep, torch.float64, L1=L1, lattice=lattice, **kw)
    loss, grad = m.loss_and_grad(torch.tensor(p))
    assert abs(loss.item() - loss0) <= 1e-9 * abs(loss0)
    # with the reference weight W(0) = 1e8 + 1 the layer gradients carry ~1e-8
    # relative rounding from either side (see test_lossv_and_reference_smoothing)
    rtol = 1e-5 if ref else 1e-8
    np.testing.assert_allclose(grad.cpu().numpy(), g0, rtol=rtol, atol=1e-12 * np.abs(g0).max())


@pytest.mark.parametrize("N,lattice,smoothing", [(16, None, None), (16, (16, 16, 16), ("gaussian", 4.0)), (12, None, ("gaussian", 4.0))])
def test_dm_kind_matches_oracle(N, lattice, smoothing):
    """dm kind (model.py:136-183, :346-347): kernel -exp(-k/kl - kh/k) k^nu,
    map = density, 5 parameters per layer, L2 by default (model.py:115-122)."""
    L, nstep = 4.0 * N, 2
    X, tgt, mk, p, _ = _problem(N, nstep, L, mask=True)
    p = p[:5 * nstep]
    orc = O.SMModelOracle(N, L, nstep, tgt, L1=False, smoothing=smoothing, mask=mk, kind="dm")
    loss0, g0 = orc.loss_and_grad(p, X)
    pm = PMesh(N, L, device=DEV, dtype=torch.float64)
    m = JERALDModel(torch.tensor(X), tgt, pm, Nstep=nstep, kind="dm", masktrain=mk, smoothing=smoothing, lattice=lattice)
    assert not m.L1
    loss, grad = m.loss_and_grad(torch.tensor(p))
    assert grad.shape == (5 * nstep,)
    assert abs(loss.item() - loss0) <= 1e-10 * abs(loss0)
    np.testing.assert_allclose(grad.cpu().numpy(), g0, rtol=1e-8, atol=1e-12 * np.abs(g0).max())
    rho = m.evaluate(torch.tensor(p))
    np.testing.assert_allclose(rho.cpu().numpy(), orc.evaluate(p, X), rtol=1e-10, atol=1e-10)


def _hi_problem(N, nstep, L, seed=83, zero_frac=0.1):
    X, tgt, mk, _, _ = _problem(N, nstep, L, seed=seed, mask=True)
    rng = np.random.default_rng(seed + 2)
    star = rng.random((N, N, N)) * 1.5
    star[rng.random((N, N, N)) < zero_frac] = 0.0  # empty cells: zero base of starmap^gamma
    lay = [[8.0, 1.2, 0.3, 1.5, 0.2, 5.0, 0.8, 0.2, 1.1, 0.3], [5.0, 0.9, 0.2, 1.6, -0.1, 3.0, 1.1, 0.25, 1.3, 0.1]]
    p = np.array(sum(lay[:nstep], []) + [1.1, 0.9, 0.3, 0.4, 0.05, 2.0, 0.3, 0.2, 1.3, 0.1, 0.02, 0.05])
    return X, tgt, mk, p, star


@pytest.mark.parametrize("N,lattice,smoothing,L1", [(16, None, None, True), (16, (16, 16, 16), ("gaussian", 4.0), False),
                                                    (12, None, ("gaussian", 4.0), True), (20, (20, 20, 20), None, False)])
def test_hi_kind_matches_oracle(N, lattice, smoothing, L1):
    """HI kind (model.py:214-247, :313-336): second source starmap^gamma2 per
    layer (10 parameters), depletion response (12), FP64 against the oracle."""
    L, nstep = 4.0 * N, 2
    X, tgt, mk, p, star = _hi_problem(N, nstep, L)
    orc = O.SMModelOracle(N, L, nstep, tgt, L1=L1, smoothing=smoothing, mask=mk, kind="HI", starmap=star)
    loss0, g0 = orc.loss_and_grad(p, X)
    pm = PMesh(N, L, device=DEV, dtype=torch.float64)
    m = JERALDModel(torch.tensor(X), tgt, pm, Nstep=nstep, kind="HI", L1=L1, masktrain=mk, smoothing=smoothing,
                    lattice=lattice, starmap=star)
    loss, grad = m.loss_and_grad(torch.tensor(p))
    assert grad.shape == (32,)
    assert abs(loss.item() - loss0) <= 1e-10 * abs(loss0)
    np.testing.assert_allclose(grad.cpu().numpy(), g0, rtol=1e-8, atol=1e-12 * np.abs(g0).max())
    F = m.evaluate(torch.tensor(p))
    np.testing.assert_allclose(F.cpu().numpy(), orc.evaluate(p, X), rtol=1e-10, atol=1e-12)


def test_hi_reference_rule_and_fp32():
    """alpha_rule reweights only alpha1 and alpha2; FP32 loss within 1e-5 of FP64."""
    N, L, nstep = 32, 128.0, 2
    X, tgt, mk, p, star = _hi_problem(N, nstep, L)
    pm = PMesh(N, L, device=DEV, dtype=torch.float64)
    out = {}
    for dt, rule in ((torch.float64, "exact"), (torch.float64, "reference"), (torch.float32, "exact")):
        m = JERALDModel(torch.tensor(X), tgt, pm, Nstep=nstep, kind="HI", L1=False, masktrain=mk, lattice=(N, N, N),
                        smoothing=("gaussian", 4.0), starmap=star, grad_rule=rule, dtype=dt)
        loss, grad = m.loss_and_grad(torch.tensor(p))
        out[(dt, rule)] = (loss.item(), grad.cpu().numpy())
    l64, g64 = out[(torch.float64, "exact")]
    _, gref = out[(torch.float64, "reference")]
    alpha = [0, 5, 10, 15]
    other = [i for i in range(32) if i not in alpha]
    np.testing.assert_allclose(gref[other], g64[other], rtol=1e-12)
    assert np.all(np.abs(gref[alpha]) < np.abs(g64[alpha]))
    orc = O.SMModelOracle(N, L, nstep, tgt, L1=False, smoothing=("gaussian", 4.0), mask=mk, kind="HI", starmap=star,
                          grad_rule="reference")
    _, g0r = orc.loss_and_grad(p, X)
    np.testing.assert_allclose(gref, g0r, rtol=1e-8, atol=1e-12 * np.abs(g0r).max())
    l32, g32 = out[(torch.float32, "exact")]
    assert abs(l32 - l64) <= 1e-5 * abs(l64)
    # per-layer and response gradients against the largest of their group
    for grp in (slice(0, 20), slice(20, 32)):
        err = np.abs(g32[grp] - g64[grp]).max() / np.abs(g64[grp]).max()
        assert err < 1e-5, (grp, err)


def _fp32_fp64_32cubed():
    N, L, nstep = 32, 128.0, 2
    X, tgt, mk, p, orc = _problem(N, nstep, L, smoothing=("gaussian", 4.0), L1=False, rms=2.0)
    out = {}
    for dt in (torch.float64, torch.float32):
        m = _model(X, tgt, mk, N, L, nstep, dt, L1=False, smoothing=("gaussian", 4.0), lattice=(N, N, N))
        loss, grad = m.loss_and_grad(torch.tensor(p))
        out[dt] = (loss.item(), grad.cpu().numpy())
    return out[torch.float64], out[torch.float32], orc.loss_and_grad(p, X)


def test_ldl_fp32_within_1e5_of_fp64():
    """32^3 lattice, 2 layers, random target: FP32 loss and every gradient but
    dL/db within 1e-5 of FP64 (dL/db: the next test)."""
    (l64, g64), (l32, g32), (loss0, g0) = _fp32_fp64_32cubed()
    assert abs(l32 - l64) <= 1e-5 * abs(l64)
    rel = np.abs(g32 - g64) / np.abs(g64)
    assert np.all(rel[:-1] < 1e-5), rel
    assert abs(l64 - loss0) <= 1e-10 * abs(loss0)


@pytest.mark.xfail(strict=True, reason="dL/db = 2 mean(F - target), and this random target's mean sits ~5% from the "
                                       "model map's: the FP32 path's mean(F), off by ~7e-7 relative, gives dL/db "
                                       "1.3e-5 from FP64 (measured); the workload configurations meet 1e-5")
def test_ldl_fp32_b_gradient_within_1e5_with_offset_target():
    (l64, g64), (l32, g32), _ = _fp32_fp64_32cubed()
    assert abs(g32[-1] - g64[-1]) <= 1e-5 * abs(g64[-1])


def _clustered_fp32_fp64():
    N, L, nstep = 16, 64.0, 2
    rng = np.random.default_rng(5)
    X = O.lattice_positions(N, L)
    X[: 3 * len(X) // 4] = 20.0 + rng.random((3 * len(X) // 4, 3)) * 6.0  # 3/4 of the particles in ~2^3 cells
    tgt = rng.random((N, N, N)) + 0.5
    p = np.array([0.8, 1.2, 0.3, 1.5, 0.2] * nstep + [1.1, 0.9, 0.05])
    out = {}
    for dt in (torch.float64, torch.float32):
        m = _model(X, tgt, None, N, L, nstep, dt, L1=False, smoothing=("gaussian", 4.0))
        loss, grad = m.loss_and_grad(torch.tensor(p))
        out[dt] = (loss.item(), grad.cpu().numpy())
    orc = O.SMModelOracle(N, L, nstep, tgt, L1=False, smoothing=("gaussian", 4.0))
    return out[torch.float64], out[torch.float32], orc.loss_and_grad(p, X)


def test_fp32_clustered_particles():
    """Most particles in a few cells (cell weight sums in the hundreds): the
    FP32 backward paint's 32-bit accumulators, scaled by the forward density's
    maximum, must not overflow; FP32 loss follows FP64, FP64 the oracle."""
    (l64, g64), (l32, g32), (loss0, g0) = _clustered_fp32_fp64()
    assert np.all(np.isfinite(g32))
    assert abs(l32 - l64) <= 1e-5 * abs(l64)
    np.testing.assert_allclose(g64, g0, rtol=1e-8, atol=1e-12 * np.abs(g0).max())


@pytest.mark.xfail(strict=True, reason="3/4 of 4096 particles within ~8 cells (densities ~400x the mean): the FP32 "
                                       "layer gradients differ from FP64 by 2-5e-5 (measured), above the 1e-5 target")
def test_fp32_clustered_particles_gradients_within_1e5():
    (l64, g64), (l32, g32), _ = _clustered_fp32_fp64()
    assert np.all(np.abs(g32 - g64) <= 1e-5 * np.abs(g64)), np.abs(g32 - g64) / np.abs(g64)


def test_sparse_particles_finite_gradients():
    """Particles in a few small clumps, most cells empty: the response's
    (rho + 1e-8)^mu sees no negative density, gradients stay finite and FP32
    follows FP64 and the oracle."""
    N, L, nstep = 32, 64.0, 2
    rng = np.random.default_rng(12)
    centres = rng.random((6, 3)) * L
    X = (centres[rng.integers(0, 6, 6000)] + rng.standard_normal((6000, 3)) * 2.0) % L
    tgt = rng.random((N, N, N)) + 0.5
    p = np.array([0.8, 1.2, 0.3, 1.5, 0.2] * nstep + [1.1, 0.9, 0.05])
    out = {}
    for dt in (torch.float64, torch.float32):
        m = _model(X, tgt, None, N, L, nstep, dt, L1=False, smoothing=("gaussian", 4.0))
        loss, grad = m.loss_and_grad(torch.tensor(p))
        out[dt] = (loss.item(), grad.cpu().numpy())
        assert np.all(np.isfinite(out[dt][1])), (dt, out[dt][1])
    (l64, g64), (l32, g32) = out[torch.float64], out[torch.float32]
    assert abs(l32 - l64) <= 1e-5 * abs(l64)
    assert np.all(np.abs(g32 - g64) <= 1e-5 * np.abs(g64)), np.abs(g32 - g64) / np.abs(g64)
    orc = O.SMModelOracle(N, L, nstep, tgt, L1=False, smoothing=("gaussian", 4.0))
    loss0, g0 = orc.loss_and_grad(p, X)
    np.testing.assert_allclose(g64, g0, rtol=1e-8, atol=1e-12 * np.abs(g0).max())


@pytest.mark.parametrize("dt", [torch.float32, torch.float64])
def test_sort_reuse_is_bit_identical(dt):
    """The backward paint on the forward's 16^3-tile sort gives the same
    integers as its own 8^3-tile sort: identical loss and gradients."""
    N, L, nstep = 32, 128.0, 2
    X, tgt, mk, p, _ = _problem(N, nstep, L, smoothing=("gaussian", 4.0), L1=False, rms=2.0)
    out = []
    for reuse in (True, False):
        m = _model(X, tgt, mk, N, L, nstep, dt, L1=False, smoothing=("gaussian", 4.0), lattice=(N, N, N),
                   reuse_sort=reuse)
        loss, grad = m.loss_and_grad(torch.tensor(p))
        out.append((loss.item(), grad.cpu().numpy()))
    assert out[0][0] == out[1][0]
    np.testing.assert_array_equal(out[0][1], out[1][1])


def test_ldl_reference_alpha_rule_and_position_grad():
    N, L, nstep = 16, 64.0, 2
    X, tgt, mk, p, _ = _problem(N, nstep, L, L1=False)
    orc = O.SMModelOracle(N, L, nstep, tgt, L1=False, grad_rule="reference")
    loss0, g0, gX0 = orc.loss_and_grad(p, X, want_pos_grad=True)
    m = _model(X, tgt, mk, N, L, nstep, torch.float64, L1=False, grad_rule="reference")
    loss, grad, gX = m.loss_and_grad_positions(torch.tensor(p))
    np.testing.assert_allclose(grad.cpu().numpy(), g0, rtol=1e-8)
    np.testing.assert_allclose(gX.cpu().numpy(), gX0, rtol=1e-7, atol=1e-12 * np.abs(gX0).max())


def test_ldl_autograd_function_and_recompute_mode():
    N, L, nstep = 16, 64.0, 2
    X, tgt, mk, p, orc = _problem(N, nstep, L, L1=False)
    _, g0 = orc.loss_and_grad(p, X)
    for last in (False, True):
        m = _model(X, tgt, mk, N, L, nstep, torch.float64, L1=False, save="positions", recompute_last=last)
        pt = torch.tensor(p, requires_grad=True)
        loss = m.loss(pt)
        loss.backward()
        np.testing.assert_allclose(pt.grad.numpy(), g0, rtol=1e-8)
    # displace/evaluate agree with the oracle
    Xf = m.displace(torch.tensor(p)).cpu().numpy()
    np.testing.assert_allclose(Xf, orc.displace(p, X), rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(m.evaluate(torch.tensor(p)).cpu().numpy(), orc.evaluate(p, X), rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("kind,N", [("sm", 32), ("dm", 32), ("sm", 30)])
def test_unmasked_l2_single_smoothing_matches_two_smoothings(kind, N):
    """Unmasked L2 with the Gaussian: sum (S r)^2 = <r, S^2 r> and the residual's
    gradient 2 S^2 r, from one smoothing with the Gaussian of width sigma sqrt(2).
    An all-ones train mask takes the path that smooths the residual and then
    its gradient; both agree to FP64 rounding (power-of-two line passes and a
    mixed-radix mesh)."""
    L, nstep = 0.8 * N, 2
    X = O.lattice_positions(N, L) + O.zeldovich_displacement(N, L, 1.5, 11)
    rng = np.random.default_rng(12)
    tgt = rng.random((N, N, N)) + 0.5
    p = np.array([0.8, 1.2, 0.3, 1.5, 0.2] * nstep + ([1.1, 0.9, 0.05] if kind == "sm" else []))
    pm = PMesh(N, L, device=DEV, dtype=torch.float64)
    res = []
    for mk in (None, np.ones((N, N, N))):
        m = JERALDModel(torch.tensor(X), tgt, pm, Nstep=nstep, kind=kind, L1=False, masktrain=mk, smoothing=("gaussian", 1.3 * L / N))
        l, g = m.loss_and_grad(torch.tensor(p))
        res.append((l.item(), g.cpu().numpy()))
    (l0, g0), (l1, g1) = res
    assert abs(l0 - l1) <= 1e-12 * abs(l1)
    np.testing.assert_allclose(g0, g1, rtol=1e-10, atol=1e-14 * np.abs(g1).max())


@pytest.mark.parametrize("dt,lattice", [(torch.float32, True), (torch.float64, False)])
def test_recomputation_keeps_the_forward_sort_bitwise(dt, lattice):
    """A recomputed layer reads the sort its forward wrote (LDL_PAINT_BINS_VALID):
    loss and gradients bitwise those of a recomputation that sorts again and of
    storing every layer."""
    N, L, nstep = 32, 25.6, 3
    X, tgt, mk, p, _ = _problem(N, nstep, L, L1=False)
    pm = PMesh(N, L, device=DEV, dtype=dt)
    res = {}
    for name, kw in (("kept", dict(save="positions", recompute_last=True)),
                     ("resorted", dict(save="positions", recompute_last=True, reuse_sort=False)),
                     ("stored", dict(save="all"))):
        m = JERALDModel(torch.tensor(X), tgt, pm, Nstep=nstep, kind="sm", L1=False, masktrain=mk,
                        lattice=(N, N, N) if lattice else None, **kw)
        l, g = m.loss_and_grad(torch.tensor(p))
        res[name] = (l.item(), g.cpu().numpy())
    for name in ("resorted", "stored"):
        assert res[name][0] == res["kept"][0]
        assert np.array_equal(res[name][1], res["kept"][1])


def test_lossv_and_reference_smoothing():
    """Reference smoothing k^-n + 1 with knorm(0) = 1e-8 (W(0) = 1e8 + 1,
    model.py:102, :376-381) and train/validation masks (model.py:430-466).
    A random train mask keeps the residual gradient non-uniform: with an
    all-ones mask the DC term makes g_rho exactly uniform and every layer
    gradient exactly zero."""
    N, L, nstep = 16, 64.0, 1
    X = O.lattice_positions(N, L) + O.zeldovich_displacement(N, L, 1.5, 5)
    rng = np.random.default_rng(6)
    tgt = rng.random((N, N, N)) + 0.5
    mt = (rng.random((N, N, N)) > 0.4).astype(float)
    mv = 1.0 - mt
    p = np.array([0.5, 0.9, 0.2, 1.3, 0.1, 1.0, 1.0, 0.0])
    orc = O.SMModelOracle(N, L, nstep, tgt, L1=True, n=1.0, mask=mt)
    orv = O.SMModelOracle(N, L, nstep, tgt, L1=True, n=1.0, mask=mv)
    pm = PMesh(N, L, device=DEV, dtype=torch.float64)
    m = JERALDModel(torch.tensor(X), tgt, pm, Nstep=nstep, kind="sm", n=1.0, masktrain=mt, maskvalid=mv)
    loss, lossv, grad = m.lossv_and_grad(torch.tensor(p))
    l0, g0 = orc.loss_and_grad(p, X)
    lv0 = orv.loss(p, X)
    assert abs(loss.item() - l0) <= 1e-9 * abs(l0)
    assert abs(lossv.item() - lv0) <= 1e-9 * abs(lv0)
    # the layer gradients come from the non-DC part of a residual whose DC
    # weight is 1e8, so they carry ~1e-8 relative rounding from either side
    np.testing.assert_allclose(grad.cpu().numpy(), g0, rtol=1e-5)


def _noncubic_multi_tile(lattice):
    nm, box, nstep = (64, 32, 128), (64.0, 40.0, 96.0), 2
    rng = np.random.default_rng(3)
    if lattice:
        lat = (32, 16, 64)
        idx = np.stack(np.meshgrid(*[np.arange(v) for v in lat], indexing="ij"), axis=-1).reshape(-1, 3)
        X = idx * (np.array(box) / np.array(lat)) + rng.uniform(-2.0, 2.0, idx.shape)
    else:
        lat = None
        X = rng.random((32768, 3)) * np.array(box)
    tgt = rng.random(nm) + 0.5
    p = np.array([0.8, 1.2, 0.3, 1.5, 0.2] * nstep + [1.1, 0.9, 0.05])
    orc = O.SMModelOracle(nm, box, nstep, tgt, L1=False, smoothing=("gaussian", 2.0))
    loss0, g0 = orc.loss_and_grad(p, X)
    out = {}
    for dt in (torch.float64, torch.float32):
        pm = PMesh(nm, box, device=DEV, dtype=dt)
        m = JERALDModel(torch.tensor(X), tgt, pm, Nstep=nstep, kind="sm", L1=False, smoothing=("gaussian", 2.0), lattice=lat)
        loss, grad = m.loss_and_grad(torch.tensor(p))
        out[dt] = (loss.item(), grad.cpu().numpy())
    return out[torch.float64], out[torch.float32], (loss0, g0)


@pytest.mark.parametrize("lattice", [False, True])
def test_ldl_noncubic_mesh_over_several_stencil_tiles(lattice):
    """A non-cubic power-of-two mesh larger than one tile of the real-space
    difference kernels along every axis (x 64 > 32-plane chunks, y 32 > 16
    rows, z 128 > 64 cells) in a non-cubic box, so halos cross tiles and
    chunks: FP64 loss and gradients against the oracle, FP32 against FP64
    (every gradient within 1e-5 for positions held as lattice displacements)."""
    (l64, g64), (l32, g32), (loss0, g0) = _noncubic_multi_tile(lattice)
    assert abs(l64 - loss0) <= 1e-10 * abs(loss0)
    np.testing.assert_allclose(g64, g0, rtol=1e-8, atol=1e-12 * np.abs(g0).max())
    assert abs(l32 - l64) <= 1e-5 * abs(l64)
    if lattice:
        assert np.all(np.abs(g32 - g64) <= 1e-5 * np.abs(g64)), np.abs(g32 - g64) / np.abs(g64)


@pytest.mark.xfail(strict=True, reason="FP32 absolute coordinates: at mesh coordinate 128 one ulp is 2^-17 cells, so CIC "
                                       "weights carry ~1e-5 errors; the nu gradients (a cancelling sum, 30x smaller than "
                                       "alpha's) differ from FP64 by 5e-5 (measured).  Lattice displacements meet 1e-5.")
def test_ldl_noncubic_absolute_fp32_gradients_within_1e5():
    (l64, g64), (l32, g32), _ = _noncubic_multi_tile(False)
    assert np.all(np.abs(g32 - g64) <= 1e-5 * np.abs(g64)), np.abs(g32 - g64) / np.abs(g64)


@pytest.mark.parametrize("kind", ["dm", "HI"])
def test_dm_and_hi_kinds_on_noncubic_multi_tile_mesh(kind):
    """The dm and HI kinds on the non-cubic mesh that spans several tiles of
    the real-space difference kernels along every axis: FP64 loss and
    gradients against the oracle."""
    nm, box, nstep = (64, 32, 128), (64.0, 40.0, 96.0), 2
    rng = np.random.default_rng(4)
    X = rng.random((32768, 3)) * np.array(box)
    tgt = rng.random(nm) + 0.5
    lay = [[8.0, 1.2, 0.3, 1.5, 0.2, 5.0, 0.8, 0.2, 1.1, 0.3], [5.0, 0.9, 0.2, 1.6, -0.1, 3.0, 1.1, 0.25, 1.3, 0.1]]
    star = None
    if kind == "dm":
        p = np.array([0.8, 1.2, 0.3, 1.5, 0.2, 0.6, 0.9, 0.2, 1.6, -0.1])
    else:
        star = rng.random(nm) * 1.5
        star[rng.random(nm) < 0.1] = 0.0
        p = np.array(lay[0] + lay[1] + [1.1, 0.9, 0.3, 0.4, 0.05, 2.0, 0.3, 0.2, 1.3, 0.1, 0.02, 0.05])
    orc = O.SMModelOracle(nm, box, nstep, tgt, L1=False, smoothing=("gaussian", 2.0), kind=kind, starmap=star)
    loss0, g0 = orc.loss_and_grad(p, X)
    pm = PMesh(nm, box, device=DEV, dtype=torch.float64)
    m = JERALDModel(torch.tensor(X), tgt, pm, Nstep=nstep, kind=kind, L1=False, smoothing=("gaussian", 2.0), starmap=star)
    loss, grad = m.loss_and_grad(torch.tensor(p))
    assert abs(loss.item() - loss0) <= 1e-10 * abs(loss0)
    np.testing.assert_allclose(grad.cpu().numpy(), g0, rtol=1e-8, atol=1e-12 * np.abs(g0).max())
