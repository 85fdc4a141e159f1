"""Helpers of the GPU tests: device propagators through the C ABI, oracle calls
on the same bytes, error norms."""
import numpy as np

from oracle import oracle as O
from plegma_amd import capi


def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        import pytest
        pytest.skip("no HIP device")
    return torch


def make_prop(L, seed, stream, t0, lam, prec=capi.F64, t_first=0, nt=None):
    """Seeded propagator on the device; returns (tensor, plg_prop view)."""
    torch = torch_cuda()
    V3 = L[0] * L[1] * L[2]
    nt = L[3] if nt is None else nt
    dt = torch.float64 if prec == capi.F64 else torch.float32
    t = torch.empty((144, nt * V3, 2), dtype=dt, device="cuda")
    v = capi.plg_prop(t.data_ptr(), prec, t_first, nt, 0, nt * V3)
    capi.synth_prop(v, seed, stream, t0, lam)
    return t, v


def cast_f32(t64, v64, scaled=False):
    """FP32 copy; with scaled=True the copy carries per-plane exponents."""
    torch = torch_cuda()
    t = torch.empty(t64.shape, dtype=torch.float32, device="cuda")
    v = capi.plg_prop(t.data_ptr(), capi.F32, v64.t_first, v64.nt, 0, v64.comp_stride)
    e = None
    if scaled:
        e = torch.empty(v64.nt, dtype=torch.int32, device="cuda")
        v.plane_exp = e.data_ptr()
    capi.cast_prop(v, v64)
    return (t, e), v


def host(t):
    import torch
    torch.cuda.synchronize()
    return np.ascontiguousarray(t.detach().to("cpu", torch.float64).numpy().reshape(-1))


def run_correlator(kind, mask, v1, v2, src, totalT, q2, mode=capi.MODE_F64,
                   space=capi.MOMENTUM_SPACE, out_prec=capi.F64):
    torch = torch_cuda()
    moms = capi.momlist(q2)
    n = capi.correlator_size(kind, mask, src, totalT, len(moms), space)
    dt = torch.float64 if out_prec == capi.F64 else torch.float32
    out = torch.full((n, 2), float("nan"), dtype=dt, device="cuda")
    capi.correlator(kind, mask, v1, v2, src, totalT, moms, mode, space, out_prec, out)
    h = host(out).reshape(n, 2)
    return h[:, 0] + 1j * h[:, 1]


def shaped(H, kind, mask, src, totalT, q2):
    ng, nd, nch = capi.kind_shape(kind, mask)
    _, _, lt = capi.window(src, totalT)
    return H.reshape(ng, nd, lt, len(capi.momlist(q2)), nch)


def rel_err(got, ref, axis=-2):
    """max over everything of |got - ref| / max_over_axis |ref| (normwise over
    momenta by default: one norm per group, dataset, row and channel)."""
    den = np.abs(ref).max(axis=axis, keepdims=True)
    den = np.where(den == 0, 1.0, den)
    return float((np.abs(got - ref) / den).max())


def oracle_H(kind, L, src, totalT, q2, P1, P2, mask=O.G_NUCL, rows=None, nthreads=0):
    return O.correlator(kind, L, src, totalT, q2, P1, P2, mask=mask, rows=rows, nthreads=nthreads)


def shell_pools(q2, minsize=24):
    """|n|^2 shells pooled in increasing order until each pool holds at least
    `minsize` momenta (a remainder below that joins the last pool): boolean
    masks over the momenta"""
    q2 = np.asarray(q2)
    pools, cur = [], np.zeros(len(q2), bool)
    for q in np.unique(q2):
        cur = cur | (q2 == q)
        if cur.sum() >= minsize:
            pools.append(cur)
            cur = np.zeros(len(q2), bool)
    if cur.any():
        if pools:
            pools[-1] = pools[-1] | cur
        else:
            pools.append(cur)
    return pools


def shell_rel_err(got, ref, moms, axis=-2):
    """error per (channel, momentum): |got - ref| / rms of |ref| over the
    momenta of similar |n|: the |n|^2 shells pooled until a pool holds at
    least 24 momenta (same group, dataset, row and channel).  The pool rms is
    the scale of a correlator at that |n|, so the n = 0 value (which dominates
    a row's maximum) does not set the scale at |n| > 0, and an accidental
    near-zero value at one momentum (n = 0 of a zero-mean channel, a shell of
    6) does not blow the error up.  Returns the max over all elements."""
    moms = np.asarray(moms).reshape(-1, 3)
    q2 = (moms ** 2).sum(1)
    got = np.moveaxis(got, axis, -1)
    ref = np.moveaxis(ref, axis, -1)
    err = np.abs(got - ref)
    den = np.empty(err.shape)
    for sel in shell_pools(q2):
        den[..., sel] = np.sqrt((np.abs(ref[..., sel]) ** 2).mean(-1, keepdims=True))
    den = np.where(den == 0, 1.0, den)
    return float((err / den).max())


def point_rel_err(got, ref):
    """pointwise |got - ref| / |ref| over all nonzero reference elements:
    (median, 99.9th percentile, max)"""
    r = np.abs(ref).reshape(-1)
    e = np.abs(got - ref).reshape(-1)
    sel = r > 0
    q = e[sel] / r[sel]
    return float(np.median(q)), float(np.quantile(q, 0.999)), float(q.max())


def host_slice(t, nt, V3, plane):
    """the 144 components of one stored time slice of a device propagator
    (tensor (144, nt*V3, 2)) as a float64 host array of an oracle view with
    comp_stride V3 and nt = 1"""
    import torch
    torch.cuda.synchronize()
    s = t.view(144, nt, V3, 2)[:, plane].to(torch.float64).cpu().numpy()
    return np.ascontiguousarray(s.reshape(-1))
