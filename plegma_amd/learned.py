"""Learned estimator of expensive correlators (host side of the learned part).

A multilayer perceptron, trained and evaluated on the GPU by the kernels of
plegma_amd/csrc/plg_mlp.hip through the C ABI (BF16 weights, FP32
accumulation), predicts a source's expensive correlators (the nucleon at
high momentum, averaged over each |n|^2 shell) from its cheap ones (the 16
mesons and the nucleon at zero momentum).  It is trained on the sources whose
correlators the plain FP64 path computed (the labelled subset); a second
labelled subset, held out of training, measures the mean error of the
predictions, which is subtracted (bias correction):
    y_i ~ f(x_i) + mean_{j in held-out} (y_j - f(x_j))
so that the mean over the predicted sources is unbiased whatever the network
learned.  Where the held-out residuals say the predictions are too poor, the
plain FP64 path computes the predicted sources instead (the fallback).

There is no counterpart of this in the reference; nothing here runs without
the HIP library.
"""
from __future__ import annotations

import numpy as np

from . import capi

# non-empty shells |n|^2 = 4..16 (no integer vector has |n|^2 = 7 or 15)
SHELLS = (4, 5, 6, 8, 9, 10, 11, 12, 13, 14, 16)
# spin projector P+ = (1 + g4)/2 of the nucleon: g4 = diag(1, 1, -1, -1) in the
# basis of lib/kernels/PLEGMA_gammas.cuh, so Tr[P+ C] = C_00 + C_11
NUCLEON_PROJ = (0 * 4 + 0, 1 * 4 + 1)
# the pion (g5, GAMMAS order 5): sum |S|^2 over the site, positive at n = 0
PION = 5


def init_params(dims, seed):
    """He-normal weights, zero biases, in the C-ABI parameter layout"""
    rng = np.random.default_rng(seed)
    parts = []
    for i, j in zip(dims[:-1], dims[1:]):
        parts.append(rng.normal(0.0, np.sqrt(2.0 / i), size=(j, i)).reshape(-1))
        parts.append(np.zeros(j))
    return np.concatenate(parts).astype(np.float32)


def _nucleon(H, moms, nrows):
    H = np.asarray(H).reshape(2, nrows, len(moms), 16)
    n2 = np.array([m[0] * m[0] + m[1] * m[1] + m[2] * m[2] for m in moms])
    return H, H[1][:, :, NUCLEON_PROJ[0]] + H[1][:, :, NUCLEON_PROJ[1]], n2       # nucleon [row][mom]


def features_targets(H, moms, nrows):
    """(x, y) of one source from its two-point result H (kind TWOPT with the
    nucleon group, momentum space) in row order: x = Re of the 16 mesons and
    the projected nucleon at n = 0 (nrows x 17), y = Re of the projected
    nucleon averaged over each shell of SHELLS (nrows x 11)."""
    H, nuc, n2 = _nucleon(H, moms, nrows)
    z = int(np.flatnonzero(n2 == 0)[0])
    x = np.concatenate([H[0][:, z, :].real, nuc[:, z:z + 1].real], axis=1)
    y = np.stack([nuc[:, n2 == s].real.mean(axis=1) for s in SHELLS], axis=1)
    return x.reshape(-1), y.reshape(-1)


def source_relative(times, t0, T):
    """row order of a source in source-relative time tau = (t - t0) mod T:
    the rows whose lattice times are `times`, sorted by tau"""
    tau = (np.asarray(times) - t0) % T
    return np.argsort(tau, kind="stable"), np.sort(tau)


def cheap_features(H0, times, src, T):
    """per time slice features from the cheap inputs (the zero-momentum
    two-point result H0: kind TWOPT with the nucleon group, one momentum
    n = 0), in source-relative time: the 16 mesons over the pion
    a(t) = Re C_pion(t, 0) and the projected nucleon over a(t)^1.5, so that
    the decay of the propagators (a power of exp(-lambda|t - t0|) per time
    slice, the same for every channel of one kind) drops out.  Returns
    x [T][17] and a [T]."""
    nrows = len(times)
    H0 = np.asarray(H0).reshape(2, nrows, 1, 16)
    order, _ = source_relative(times, src[3], T)
    a = H0[0][:, 0, PION].real
    if np.any(a <= 0):
        raise ValueError("the zero-momentum pion must be positive")
    nuc = (H0[1][:, 0, NUCLEON_PROJ[0]] + H0[1][:, 0, NUCLEON_PROJ[1]]).real
    x = np.concatenate([H0[0][:, 0, :].real / a[:, None], (nuc / a ** 1.5)[:, None]], axis=1)
    return x[order], a[order]


def shell_targets(H, moms, times, src, L):
    """the expensive outputs of a source in source-relative time: Re of the
    projected nucleon with the source phase removed, exp(-i 2 pi n.x0 / L) C(t, n)
    (the momentum phases then refer to the lattice origin, as for every
    source), averaged over each |n|^2 shell of SHELLS: y [T][11]"""
    T = L[3]
    nrows = len(times)
    _, nuc, n2 = _nucleon(H, moms, nrows)
    m = np.asarray(moms, dtype=np.int64).reshape(-1, 3)
    ph = np.exp(-2j * np.pi * (m * (np.asarray(src[:3]) / np.asarray(L[:3]))).sum(1))
    nuc = nuc * ph[None, :]
    y = np.stack([nuc[:, n2 == s].real.mean(axis=1) for s in SHELLS], axis=1)
    order, _ = source_relative(times, src[3], T)
    return y[order]


class Standardizer:
    """per-feature affine map to zero mean and unit variance over a sample"""

    def __init__(self, a):
        a = np.asarray(a, dtype=np.float64)
        self.mu = a.mean(axis=0)
        self.sd = a.std(axis=0)
        self.sd[self.sd == 0] = 1.0

    def __call__(self, a):
        return (np.asarray(a, dtype=np.float64) - self.mu) / self.sd

    def inverse(self, z):
        return np.asarray(z, dtype=np.float64) * self.sd + self.mu


class CorrelatorNet:
    """the network with its input/output standardisation"""

    def __init__(self, n_in, n_out, hidden=(256, 256), seed=0, max_batch=256):
        import torch
        self.torch = torch
        self.dims = [int(n_in), *[int(h) for h in hidden], int(n_out)]
        self.max_batch = max_batch
        self.net = capi.Mlp(self.dims, max_batch)
        self.net.set_params(init_params(self.dims, seed))
        self.sx = self.sy = None
        self.history = []

    def _dev(self, a):
        return self.torch.tensor(np.ascontiguousarray(a, dtype=np.float32), device="cuda")

    def fit(self, x, y, steps=2000, lr=1e-3, chunk=100):
        """full-batch Adam on the standardised pairs; loss history per chunk"""
        self.sx, self.sy = Standardizer(x), Standardizer(y)
        xd, yd = self._dev(self.sx(x)), self._dev(self.sy(y))
        done = 0
        while done < steps:
            k = min(chunk, steps - done)
            self.history.append((done + k, self.net.train(xd, yd, k, lr)))
            done += k
        return self.history[-1][1] if self.history else None

    def predict(self, x):
        x = np.asarray(x, dtype=np.float64)
        out = []
        for i in range(0, len(x), self.max_batch):
            xd = self._dev(self.sx(x[i:i + self.max_batch]))
            yd = self.torch.empty((xd.shape[0], self.dims[-1]), dtype=self.torch.float32, device="cuda")
            self.net.forward(xd, yd)
            self.torch.cuda.synchronize()
            out.append(self.sy.inverse(yd.cpu().numpy()))
        return np.concatenate(out, axis=0)

    def close(self):
        self.net.close()


class Estimate:
    """bias-corrected predictions of the unlabelled sources"""

    def __init__(self, pred, delta, resid_rms, scale):
        self.pred = pred              # [n_pred][n_out] corrected predictions
        self.delta = delta            # [n_out] mean held-out residual (subtracted bias)
        self.resid_rms = resid_rms    # [n_out] rms of held-out residuals after correction
        self.scale = scale            # [n_out] rms of held-out targets
        self.used_fallback = False

    @property
    def mean(self):
        return self.pred.mean(axis=0)

    def rel_error_estimate(self):
        """held-out residual rms relative to the held-out signal, per output"""
        return self.resid_rms / np.where(self.scale == 0, 1.0, self.scale)


class RowwiseNet:
    """the network applied to each time slice of a source (weights shared
    over the T time slices): a source's inputs are its T x 17 features and
    the T pion values a(t); its outputs the T x 11 shell values, the network
    predicting them in units of a(t)^1.5 (the decay-free ratios the features
    are).  fit() takes the time slices of the training sources as samples."""

    def __init__(self, T, hidden=(256, 256), seed=0, max_batch=4096):
        self.T = T
        self.net = CorrelatorNet(17, len(SHELLS), hidden=hidden, seed=seed, max_batch=max_batch)

    @staticmethod
    def pack(x, a):
        """one source's network input: x [T][17] and a [T] in one row"""
        return np.concatenate([np.asarray(x).reshape(-1), np.asarray(a).reshape(-1)])

    def _split(self, X):
        X = np.asarray(X, dtype=np.float64)
        x = X[:, :17 * self.T].reshape(-1, 17)
        a = X[:, 17 * self.T:].reshape(-1)
        return x, a

    def fit(self, X, Y, steps=2000, lr=1e-3):
        x, a = self._split(X)
        z = np.asarray(Y, dtype=np.float64).reshape(-1, len(SHELLS)) / a[:, None] ** 1.5
        return self.net.fit(x, z, steps=steps, lr=lr)

    def predict(self, X):
        x, a = self._split(X)
        y = self.net.predict(x) * a[:, None] ** 1.5
        return y.reshape(len(X), self.T * len(SHELLS))

    @property
    def history(self):
        return self.net.history

    def close(self):
        self.net.close()


def bias_corrected(net, x_bc, y_bc, x_pred):
    f_bc = net.predict(x_bc)
    y_bc = np.asarray(y_bc, dtype=np.float64)
    delta = (y_bc - f_bc).mean(axis=0)
    resid = y_bc - f_bc - delta
    return Estimate(net.predict(x_pred) + delta, delta, np.sqrt((resid ** 2).mean(axis=0)),
                    np.sqrt((y_bc ** 2).mean(axis=0)))


def expected_gain(est, y_lab, n_lab, n_bc, n_pred):
    """the mean-squared-error ratio of the two estimators of the mean over all
    sources, labelled-only over bias-corrected, estimated from the labelled
    and held-out sources alone.  With s_y the spread of the outputs over
    sources and s_r that of the corrected residuals, the two errors are
    (n_pred/N)^2 s_y^2 (1/n_lab + 1/n_pred) and (n_pred/N)^2 s_r^2 (1/n_bc +
    1/n_pred); outputs are weighted by their spread s_y (relative errors)."""
    y_lab = np.asarray(y_lab, dtype=np.float64)
    sy2 = y_lab.var(axis=0, ddof=1)
    w = np.where(sy2 > 0, 1.0 / np.where(sy2 > 0, sy2, 1.0), 0.0)
    lab = np.sum(w * sy2) * (1.0 / n_lab + 1.0 / n_pred)
    ml = np.sum(w * est.resid_rms ** 2) * (1.0 / n_bc + 1.0 / n_pred)
    return float(lab / ml) if ml > 0 else float("inf")


def with_fallback(est, exact, tol=None, gain=None, min_gain=1.0):
    """replace the predictions by exact(), the plain FP64 path over the
    predicted sources, when the held-out relative error estimate exceeds tol
    for any output (tol given), or when the expected gain of the estimator
    (expected_gain) is below min_gain (gain given)"""
    bad = (tol is not None and np.max(est.rel_error_estimate()) > tol) or \
          (gain is not None and gain < min_gain)
    if bad:
        est.pred = np.asarray(exact(), dtype=np.float64)
        est.used_fallback = True
    return est
