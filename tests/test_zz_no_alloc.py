"""plg_correlator_static (no device allocation, static scratch) against
plg_correlator on the same inputs: two calls in a row, the second on a smaller
correlator (shorter window, fewer momenta, another kind), FP64 and mixed; and
the refusal of a correlator above the static maximum before any work is
queued.  Tolerances as the parity tests: FP64 1e-12 and mixed 1e-6, normwise
over momenta (the two entry points may use different K splits of the
projection, so the FP64 sums need not agree to the last bit)."""
import numpy as np
import pytest

from plegma_amd import capi

import gpu_util as G

pytestmark = pytest.mark.gpu

L = [8, 8, 8, 16]
MASK = capi.G_NUCL | capi.G_DELTAPP_ALL


def both(kind, mask, v1, v2, src, totalT, q2, mode):
    torch = G.torch_cuda()
    moms = capi.momlist(q2)
    n = capi.correlator_size(kind, mask, src, totalT, len(moms), capi.MOMENTUM_SPACE)
    outs = []
    for fn in (capi.correlator, capi.correlator_static):
        out = torch.full((n, 2), float("nan"), dtype=torch.float64, device="cuda")
        fn(kind, mask, v1, v2, src, totalT, moms, mode, capi.MOMENTUM_SPACE, capi.F64, out)
        h = G.host(out).reshape(n, 2)
        outs.append(G.shaped(h[:, 0] + 1j * h[:, 1], kind, mask, src, totalT, q2))
    return outs


def test_static_matches_correlator_twice():
    G.torch_cuda()
    print("\n[gpu] no-alloc: plg_init", L, flush=True)
    capi.init(L, 0)
    try:
        src = [3, 5, 1, 14]
        t64, v64 = G.make_prop(L, 77, 3, src[3], 0.5)
        (t32, e32), v32 = G.cast_f32(t64, v64, scaled=True)
        # first call: full window, 16 mesons + proton + Delta++, |n|^2 <= 4
        for mode, v, tol in ((capi.MODE_F64, v64, 1e-12), (capi.MODE_MIXED, v32, 1e-6)):
            print(f"[gpu] no-alloc: twopt mode {mode}, totalT 16, q2 4", flush=True)
            ref, got = both(capi.TWOPT, MASK, v, v, src, 16, 4, mode)
            assert np.all(np.isfinite(got))
            e = G.rel_err(got, ref)
            print(f"[gpu] no-alloc: twopt mode {mode} rel err {e:.2e}", flush=True)
            assert e < tol, (mode, e)
        # second call, smaller: a wrapped window of 6 slices, |n|^2 <= 2, the
        # 10 meson channels with two datasets
        for mode, v, tol in ((capi.MODE_F64, v64, 1e-12), (capi.MODE_MIXED, v32, 1e-6)):
            print(f"[gpu] no-alloc: mesons mode {mode}, totalT 6, q2 2", flush=True)
            ref, got = both(capi.MESONS, 0, v, v, src, 6, 2, mode)
            assert np.all(np.isfinite(got))
            e = G.rel_err(got, ref)
            print(f"[gpu] no-alloc: mesons mode {mode} rel err {e:.2e}", flush=True)
            assert e < tol, (mode, e)
    finally:
        capi.finalize()


def test_static_refuses_above_maximum():
    """16^3 x 32, FP64, 80 channels: the site correlator alone is
    32 x 80 x 4096 x 16 B = 168 MB, above the 128 MiB of static scratch"""
    torch = G.torch_cuda()
    L2 = [16, 16, 16, 32]
    print("\n[gpu] no-alloc: plg_init", L2, flush=True)
    capi.init(L2, 0)
    try:
        src = [0, 0, 0, 0]
        t64, v64 = G.make_prop(L2, 5, 0, 0, 0.5)
        moms = capi.momlist(1)
        n = capi.correlator_size(capi.TWOPT, MASK, src, 32, len(moms), capi.MOMENTUM_SPACE)
        out = torch.zeros((n, 2), dtype=torch.float64, device="cuda")
        with pytest.raises(capi.PlgError) as e:
            capi.correlator_static(capi.TWOPT, MASK, v64, v64, src, 32, moms, capi.MODE_F64,
                                   capi.MOMENTUM_SPACE, capi.F64, out)
        assert e.value.code == -6
        assert "static maximum" in str(e.value)
        torch.cuda.synchronize()
        assert float(out.abs().max()) == 0.0   # nothing was written
        # a correlator that fits still runs after the refusal
        ref, got = both(capi.MESONS16, 0, v64, v64, src, 4, 1, capi.MODE_F64)
        assert G.rel_err(got, ref) < 1e-12
    finally:
        capi.finalize()
