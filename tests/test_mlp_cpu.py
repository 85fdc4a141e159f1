"""CPU checks of the learned part: the network oracle against finite
differences and a closed form, its bf16 rounding against torch's, and the
host logic of the bias-corrected estimator (no GPU needed)."""
import numpy as np
import pytest

from oracle import mlp as M
from plegma_amd import learned


def test_bf16_rounding_matches_torch():
    import torch
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.normal(size=2000) * 10.0 ** rng.integers(-8, 8, size=2000),
                        # exact ties: 1 + k 2^-8 + 2^-9 round to even
                        1.0 + np.arange(16) * 2.0 ** -7 + 2.0 ** -8, [0.0, -0.0, 1.0, -3.5]])
    ref = torch.tensor(x.astype(np.float32)).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(M.bf16(x), ref)


def test_gradient_matches_finite_differences():
    dims = [5, 7, 6, 3]
    rng = np.random.default_rng(1)
    p = rng.normal(size=M.nparams(dims)) * 0.5
    x = rng.normal(size=(4, 5))
    t = rng.normal(size=(4, 3))
    _, g = M.loss_grad(p, dims, x, t, emulate=False)
    h = 1e-6
    fd = np.empty_like(p)
    for i in range(len(p)):
        e = np.zeros_like(p)
        e[i] = h
        fd[i] = (M.loss_grad(p + e, dims, x, t, False)[0] - M.loss_grad(p - e, dims, x, t, False)[0]) / (2 * h)
    assert np.abs(g - fd).max() <= 1e-7 * max(1.0, np.abs(fd).max())


def test_single_layer_is_least_squares():
    """one Linear layer: loss and gradient of linear least squares in closed form"""
    dims = [4, 2]
    rng = np.random.default_rng(2)
    W, b = rng.normal(size=(2, 4)), rng.normal(size=2)
    p = np.concatenate([W.reshape(-1), b])
    x, t = rng.normal(size=(6, 4)), rng.normal(size=(6, 2))
    loss, g = M.loss_grad(p, dims, x, t, emulate=False)
    r = x @ W.T + b - t
    assert loss == pytest.approx((r ** 2).mean(), rel=1e-14)
    gW = 2.0 / r.size * r.T @ x
    gb = 2.0 / r.size * r.sum(axis=0)
    np.testing.assert_allclose(g, np.concatenate([gW.reshape(-1), gb]), rtol=1e-13, atol=1e-15)


def test_emulated_rounding_stays_close_to_fp64():
    dims = [16, 32, 8]
    rng = np.random.default_rng(4)
    p = learned.init_params(dims, 5).astype(np.float64)
    x, t = rng.normal(size=(8, 16)), rng.normal(size=(8, 8))
    l64, g64 = M.loss_grad(p, dims, x, t, emulate=False)
    lbf, gbf = M.loss_grad(p, dims, x, t, emulate=True)
    assert abs(lbf - l64) <= 2e-2 * l64
    assert np.linalg.norm(gbf - g64) <= 3e-2 * np.linalg.norm(g64)


def test_param_layout():
    dims = [1088, 256, 256, 704]
    p = learned.init_params(dims, 0)
    assert p.size == M.nparams(dims) == 525504
    (W0, b0), (W1, b1), (W2, b2) = M.layers(p.astype(np.float64), dims)
    assert W0.shape == (256, 1088) and b2.shape == (704,)
    assert np.all(b0 == 0) and abs(W0.std() - np.sqrt(2 / 1088)) < 0.01 * np.sqrt(2 / 1088) * 10


class _Fixed:
    """a predictor with a known error: f(x) = g(x) + bias"""

    def __init__(self, g, bias):
        self.g, self.bias = g, bias

    def predict(self, x):
        return self.g(x) + self.bias


def test_bias_correction_is_unbiased():
    """the corrected mean over predicted sources is unbiased even for a biased
    predictor; the uncorrected one is not"""
    rng = np.random.default_rng(7)
    a = rng.normal(size=6)
    g = lambda x: np.tanh(x @ a)[:, None]
    est_c, est_u, true = [], [], []
    for _ in range(400):
        x = rng.normal(size=(96, 6))
        y = g(x) + 0.3 * rng.normal(size=(96, 1))
        f = _Fixed(g, 0.25)
        e = learned.bias_corrected(f, x[:32], y[:32], x[32:])
        est_c.append(e.mean[0])
        est_u.append(f.predict(x[32:]).mean())
        true.append(y[32:].mean())
        corr, mean = M.bias_corrected(f.predict(x[32:]), f.predict(x[:32]), y[:32])
        assert np.allclose(corr, e.pred) and np.allclose(mean, e.mean)
    d = np.array(est_c) - np.array(true)
    assert abs(d.mean()) < 3 * d.std() / np.sqrt(len(d))
    assert abs(np.mean(np.array(est_u) - np.array(true)) - 0.25) < 0.02


def test_fallback_switches_to_exact():
    rng = np.random.default_rng(8)
    x = rng.normal(size=(40, 3))
    y = rng.normal(size=(40, 2))
    e = learned.bias_corrected(_Fixed(lambda v: np.zeros((len(v), 2)), 0.0), x[:10], y[:10], x[10:])
    exact = lambda: y[10:]
    e2 = learned.with_fallback(e, exact, tol=0.5)     # noise-only predictor: error ~ 1
    assert e2.used_fallback and np.array_equal(e2.pred, y[10:])
    e3 = learned.bias_corrected(_Fixed(lambda v: np.zeros((len(v), 2)), 0.0), x[:10], y[:10], x[10:])
    assert not learned.with_fallback(e3, exact, tol=10.0).used_fallback


def test_features_targets_layout():
    from plegma_amd import capi
    moms = capi.momlist(16)
    nrows = 3
    rng = np.random.default_rng(9)
    H = rng.normal(size=(2, nrows, len(moms), 16)) + 1j * rng.normal(size=(2, nrows, len(moms), 16))
    x, y = learned.features_targets(H.reshape(-1), moms, nrows)
    assert x.shape == (nrows * 17,) and y.shape == (nrows * 11,)
    x = x.reshape(nrows, 17)
    y = y.reshape(nrows, 11)
    assert np.array_equal(x[:, :16], H[0, :, 0, :].real)                 # n = 0 is the first momentum
    assert np.allclose(x[:, 16], (H[1, :, 0, 0] + H[1, :, 0, 5]).real)
    sel = [i for i, m in enumerate(moms) if m[0] ** 2 + m[1] ** 2 + m[2] ** 2 == 9]
    assert len(sel) == 30
    assert np.allclose(y[:, learned.SHELLS.index(9)], (H[1][:, sel, 0] + H[1][:, sel, 5]).real.mean(1))


def test_source_relative_features_drop_the_decay():
    """cheap_features: rows go to source-relative time, and per-time-slice
    decay factors (d^2 on the mesons, d^3 on the nucleon) cancel"""
    T, t0 = 8, 5
    rng = np.random.default_rng(10)
    times = list(range(T))
    H0 = rng.normal(size=(2, T, 1, 16)) + 1j * rng.normal(size=(2, T, 1, 16))
    H0[0, :, 0, learned.PION] = np.abs(H0[0, :, 0, learned.PION]) + 1.0
    x, a = learned.cheap_features(H0.reshape(-1), times, [0, 0, 0, t0], T)
    order = [(t0 + k) % T for k in range(T)]
    np.testing.assert_allclose(a, H0[0, order, 0, learned.PION].real)
    np.testing.assert_allclose(x[:, learned.PION], 1.0)
    d = np.exp(-0.7 * np.abs(np.arange(T) - t0))
    Hd = H0.copy()
    Hd[0] *= (d ** 2)[:, None, None]
    Hd[1] *= (d ** 3)[:, None, None]
    xd, ad = learned.cheap_features(Hd.reshape(-1), times, [0, 0, 0, t0], T)
    np.testing.assert_allclose(xd, x, rtol=1e-12)
    np.testing.assert_allclose(ad, a * (d ** 2)[order], rtol=1e-12)


def test_shell_targets_remove_the_source_phase():
    """C(t, n) = exp(+i 2pi n.x0/L) D(t, n) for a source at x0: the targets
    are the shell means of Re D, the same for every x0"""
    from plegma_amd import capi
    L = [4, 6, 8, 4]
    moms = capi.momlist(16)
    m = np.array(moms)
    rng = np.random.default_rng(11)
    D = rng.normal(size=(2, L[3], len(moms), 16)) + 1j * rng.normal(size=(2, L[3], len(moms), 16))
    times = list(range(L[3]))
    ys = []
    for x0 in ([0, 0, 0], [1, 5, 3], [3, 2, 7]):
        ph = np.exp(2j * np.pi * (m * (np.array(x0) / np.array(L[:3]))).sum(1))
        C = D * ph[None, None, :, None]
        ys.append(learned.shell_targets(C.reshape(-1), moms, times, x0 + [0], L))
    for y in ys[1:]:
        np.testing.assert_allclose(y, ys[0], rtol=1e-12, atol=1e-12)
    nuc = D[1][:, :, 0] + D[1][:, :, 5]
    sel = (m ** 2).sum(1) == 9
    np.testing.assert_allclose(ys[0][:, learned.SHELLS.index(9)], nuc[:, sel].real.mean(1))


def test_expected_gain_matches_monte_carlo():
    """expected_gain from one draw's labelled and held-out sources agrees with
    the mean-squared-error ratio of the two estimators over many draws"""
    rng = np.random.default_rng(12)
    nlab, nbc, npred = 64, 32, 192
    N = nlab + npred
    sy, sr = 1.0, 0.3
    el, em, gains = [], [], []
    for _ in range(300):
        u = rng.normal(size=(N, 3)) * sy                 # outputs of N sources
        pred = u + rng.normal(size=(N, 3)) * sr + 0.2    # a biased predictor
        est = learned.bias_corrected(_Fixed(lambda v: v, 0.0), pred[nlab - nbc:nlab], u[nlab - nbc:nlab],
                                     pred[nlab:])
        truth = u.mean(0)
        m_ml = (u[:nlab].sum(0) + est.pred.sum(0)) / N
        m_lab = u[:nlab].mean(0)
        el.append(((m_lab - truth) ** 2).sum())
        em.append(((m_ml - truth) ** 2).sum())
        gains.append(learned.expected_gain(est, u[:nlab], nlab, nbc, npred))
    mc = np.mean(el) / np.mean(em)
    assert abs(np.median(gains) / mc - 1) < 0.25, (np.median(gains), mc)
    assert mc > 3          # sr = 0.3 sy: the corrected estimator wins
