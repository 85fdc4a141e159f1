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


def init_params(dims, seed):
    """He-normal weights, zero biases, in the C-ABI parameter layout"""
    rng = np.random.default_rng(seed)
    parts = []
    for i, j in zip(dims[:-1], dims[1:]):
        parts.append(rng.normal(0.0, np.sqrt(2.0 / i), size=(j, i)).reshape(-1))
        parts.append(np.zeros(j))
    return np.concatenate(parts).astype(np.float32)


def features_targets(H, moms, nrows):
    """(x, y) of one source from its two-point result H (kind TWOPT with the
    nucleon group, momentum space): x = Re of the 16 mesons and the projected
    nucleon at n = 0 over the rows (nrows x 17), y = Re of the projected
    nucleon averaged over each shell of SHELLS (nrows x 11)."""
    H = np.asarray(H).reshape(2, nrows, len(moms), 16)
    n2 = np.array([m[0] * m[0] + m[1] * m[1] + m[2] * m[2] for m in moms])
    z = int(np.flatnonzero(n2 == 0)[0])
    nuc = H[1][:, :, NUCLEON_PROJ[0]] + H[1][:, :, NUCLEON_PROJ[1]]     # [row][mom]
    x = np.concatenate([H[0][:, z, :].real, nuc[:, z:z + 1].real], axis=1)
    y = np.stack([nuc[:, n2 == s].real.mean(axis=1) for s in SHELLS], axis=1)
    return x.reshape(-1), y.reshape(-1)


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


def bias_corrected(net, x_bc, y_bc, x_pred):
    f_bc = net.predict(x_bc)
    y_bc = np.asarray(y_bc, dtype=np.float64)
    delta = (y_bc - f_bc).mean(axis=0)
    resid = y_bc - f_bc - delta
    return Estimate(net.predict(x_pred) + delta, delta, np.sqrt((resid ** 2).mean(axis=0)),
                    np.sqrt((y_bc ** 2).mean(axis=0)))


def with_fallback(est, exact, tol):
    """replace the predictions by exact(), the plain FP64 path over the
    predicted sources, when the held-out relative error estimate exceeds tol
    for any output"""
    if np.max(est.rel_error_estimate()) > tol:
        est.pred = np.asarray(exact(), dtype=np.float64)
        est.used_fallback = True
    return est
