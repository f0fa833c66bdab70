"""Training pieces around the path (model.py:492-659, JERALD/runJERALD.py:114-221):
optimizers, state files, masks, bounds, schedule and the loop (CPU), and a
few optimizer steps of the GPU model."""
import numpy as np
import pytest

from jerald_amd import train as T


def test_adam_and_rmsprop_first_step_is_sign_step():
    g = np.array([3.0, -0.5, 1e-3])
    p0 = np.array([1.0, 2.0, 3.0])
    a = T.Adam(LR=0.1, eps=0.0)
    a.init(p0)
    # bias-corrected first moments: mu_hat = g, nu_hat = g^2
    np.testing.assert_allclose(a.step(p0, g), p0 - 0.1 * np.sign(g), rtol=1e-12)
    r = T.RMSProp(LR=0.1, eps=0.0)
    r.init(p0)
    np.testing.assert_allclose(r.step(p0, g), p0 - 0.1 * np.sign(g), rtol=1e-12)


@pytest.mark.parametrize("opt", [T.Adam(LR=0.05), T.RMSProp(LR=0.02)])
def test_optimizers_minimise_a_quadratic(opt):
    a = np.array([1.0, 10.0, 0.1])
    c = np.array([0.5, -2.0, 3.0])
    p = np.zeros(3)
    opt.init(p)
    for _ in range(3000):
        p = opt.step(p, 2 * a * (p - c))
    np.testing.assert_allclose(p, c, atol=0.05)


def test_adam_state_file_round_trip(tmp_path):
    a = T.Adam(LR=lambda i: T.lr_schedule(i, n=3))
    a.init(np.zeros(3))
    p = np.array([1.0, -2.0, 0.25])
    for g in ([1.0, 2.0, 3.0], [-1.0, 0.5, 2.0]):
        p = a.step(p, np.array(g))
    f = tmp_path / "opt.txt"
    a.save_state(f, p)
    lines = f.read_text().split("\n")
    assert lines[0] == "mu" and lines[4] == "nu" and lines[8] == "itr" and lines[9] == "2" and lines[10] == "LR"
    b = T.Adam()
    found, pb = b.import_state(f, paramsreq=True)
    assert found
    np.testing.assert_array_equal(pb, p)
    np.testing.assert_array_equal(b.mu, a.mu)
    np.testing.assert_array_equal(b.nu, a.nu)
    assert b.itr == 2
    # the continued optimiser takes the same next step
    g = np.array([0.3, -0.1, 0.7])
    b.LR = a.LR
    np.testing.assert_array_equal(b.step(pb, g), a.step(p, g))


def test_state_file_text_is_the_savetxt_layout(tmp_path):
    """The writer's text is the layout model.py:636-659 produces with
    numpy.savetxt (%.18e per value, one per line) for array learning rates."""
    import io
    a = T.Adam(LR=lambda i: T.lr_schedule(i, n=2))
    a.init(np.zeros(2))
    p = a.step(np.array([0.5, -1.5]), np.array([0.2, -0.4]))
    f = tmp_path / "s.txt"
    a.save_state(f, p)
    buf = io.StringIO()
    for key, vals in (("mu", a.mu), ("nu", a.nu)):
        buf.write(key + "\n")
        np.savetxt(buf, vals)
    buf.write("itr\n1\nLR\n")
    np.savetxt(buf, a.LR(1))
    buf.write("params\n")
    np.savetxt(buf, p)
    assert f.read_text() == buf.getvalue()


def test_adam_import_without_params_warns(tmp_path):
    f = tmp_path / "opt.txt"
    f.write_text("mu\n1.0\n2.0\nnu\n3.0\n4.0\nitr\n7\nLR\n0.01\n")
    a = T.Adam()
    with pytest.warns(UserWarning):
        assert a.import_state(f, paramsreq=True) == (False, 0)
    np.testing.assert_array_equal(a.mu, [1.0, 2.0])
    assert a.itr == 7


def test_lr_schedule():
    lr = [float(T.lr_schedule(i)[0]) for i in (0, 10, 100, 1000, 10 ** 5)]
    assert lr[0] == pytest.approx(0.002)
    assert lr[-1] == pytest.approx(0.02)
    assert all(x < y for x, y in zip(lr, lr[1:]))
    # closed form at i = tau: (LRf - LRi) tanh(1/2) + LRi
    assert lr[2] == pytest.approx(0.018 * np.tanh(0.5) + 0.002)


@pytest.mark.parametrize("N", [16, 20, 64])
def test_split_masks(N):
    tr, va, te = T.split_masks(N)
    assert np.all(tr + va + te <= 1)  # disjoint
    nv, nt = int(0.5 * N), int(0.43 * N)
    assert tr.sum() == N * N * (N - nv)
    assert te.sum() == nt ** 3
    assert va.sum() == N * N * nv - nt ** 3
    i = np.arange(N)
    assert np.all(tr[:, :, i < nv] == 0) and np.all(tr[:, :, i >= nv] == 1)


def test_initial_params_bounds_and_clip():
    for tgt, n0, nb in [("dm", 10, 10), ("Mstar", 14, 13), ("HI", 32, 32)]:
        assert len(T.initial_params(tgt, 2)) == n0
        assert len(T.bounds(tgt, 2)) == nb
    b = T.bounds("Mstar", 1)
    p = T.clip([1, 1, -3, 0.0, 1, 0.0, 0.0, 5, 9], b)
    np.testing.assert_array_equal(p, [1, 1, 0.0, 0.1, 1, 0.1, 1e-5, 5, 9])


def test_prepare_starmap_and_hi_target_scale():
    s = T.prepare_starmap(np.array([[0.0, -1.0], [2.0, 6.0]]))
    floor = 1e-12 / ((2e-12 + 8.0) / 4)
    np.testing.assert_allclose(s, [[floor, floor], [2.0 / 2.0000000000005, 6.0 / 2.0000000000005]], rtol=1e-12)
    assert s.mean() == pytest.approx(1.0)
    assert T.hi_target_scale(2) == (4e-11, -1e-3)
    with pytest.raises(ValueError):
        T.hi_target_scale(1)


class _Quadratic:
    """Stand-in model: train loss |p - c|^2, validation loss |p - c'|^2."""

    def __init__(self, c, cv):
        self.c, self.cv = np.asarray(c), np.asarray(cv)

    def lossv_and_grad(self, p):
        return np.sum((p - self.c) ** 2), np.sum((p - self.cv) ** 2), 2 * (p - self.c)


def test_train_loop_tracks_best_validation_and_writes_files(tmp_path):
    m = _Quadratic([1.0, 0.2, 3.0], [1.0, 0.2, 2.5])
    bnds = [(None, None), (0.5, None), (None, None)]  # c[1] lies below its bound
    opt = T.Adam(LR=0.1)
    p0 = np.array([0.0, 1.0, 0.0])
    opt.init(p0)
    best, bv, step, hist = T.train(m, p0, bnds, Nopt=60, optimizer=opt, paramsave=tmp_path / "best.txt",
                                   optf=tmp_path / "opt.txt", log=lambda s: None)
    assert bv == min(v for _, v in hist)
    assert hist[step - 1][1] == bv
    assert best[1] >= 0.5  # clipped every step
    assert hist[-1][0] < hist[0][0]
    np.testing.assert_array_equal(np.loadtxt(tmp_path / "best.txt"), best)
    assert (tmp_path / "opt.txt").read_text().startswith("mu\n")


@pytest.mark.gpu
def test_training_steps_on_gpu_reduce_loss():
    import torch

    from jerald_amd.model import JERALDModel
    from jerald_amd.pmesh import PMesh
    from oracle import ldl_oracle as O
    N, L, nstep = 16, 64.0, 1
    X = O.lattice_positions(N, L) + O.zeldovich_displacement(N, L, 1.5, 3)
    pm = PMesh(N, L, device="cuda", dtype=torch.float64)
    p_true = np.array([1.0, 0.5, 1.0, 8.0, 0.0, 1.0, 1.0, 0.0])
    gen = JERALDModel(torch.tensor(X), None, pm, Nstep=nstep)
    target = gen.evaluate(torch.tensor(p_true))
    tr, va, _ = T.split_masks(N)
    model = JERALDModel(torch.tensor(X), target, pm, Nstep=nstep, masktrain=tr, maskvalid=va)
    # 5 Nstep + 4 entries, as the reference passes for 'Mstar'; response started away from the truth
    p0 = np.array([1.0, 0.5, 1.0, 8.0, 0.0, 1.4, 0.6, 0.2, 0.0])
    opt = T.Adam(LR=0.05)
    opt.init(p0)
    best, bv, step, hist = T.train(model, p0, T.bounds("Mstar", nstep), Nopt=8, optimizer=opt, log=lambda s: None)
    assert hist[-1][0] < 0.75 * hist[0][0]  # 8 Adam steps of 0.05: 0.27 -> 0.15 here
    assert bv < hist[0][1]
