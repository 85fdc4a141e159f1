"""GPU parity at the largest single-GPU workload, against the CPU oracle.

Lattice 64^3 x 128 (33.5 M sites), one source at the origin, two flavour
propagators stored in FP32 with per-time-slice exponents (light, decay 0.5;
heavy, decay 1.2: 2 x 38.65 GB), 16 heavy-light mesons + 4 meson-meson
channels, |n|^2 <= 9 (123 momenta), all 128 time slices, through
plg_correlator (mixed path: FP32 contraction, split-bf16 MFMA projection, FP64
time-slice sums).  The oracle recomputes sampled time slices in FP64 from the
FP64 propagator values the FP32 data were rounded from.  Tolerances: normwise
over momenta 1e-6; per (channel, momentum) against the rms of the reference
over the momenta of similar |n| (|n|^2 shells pooled to at least 24
momenta) 5e-6.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from plegma_amd import capi

import gpu_util as G

pytestmark = pytest.mark.gpu

L = [64, 64, 64, 128]
V3 = L[0] * L[1] * L[2]
Q = 9
SEED = 20261202
LAMBDAS = (0.5, 1.2)
NTH = min(os.cpu_count() or 1, 16)


@pytest.fixture(scope="module", autouse=True)
def ctx():
    G.torch_cuda()
    print("\n[gpu] plg_init", L, flush=True)
    capi.init(L, 0)
    yield
    capi.finalize()


def test_largest_workload():
    torch = G.torch_cuda()
    T = L[3]
    src = [0, 0, 0, 0]
    keep, views = [], []
    for fl, lam in enumerate(LAMBDAS):
        print(f"[gpu] largest: FP32 flavour {fl} (38.65 GB)", flush=True)
        t = torch.empty((144, T * V3, 2), dtype=torch.float32, device="cuda")
        e = torch.empty(T, dtype=torch.int32, device="cuda")
        v = capi.plg_prop(t.data_ptr(), capi.F32, 0, T, 0, T * V3, e.data_ptr())
        capi.synth_prop(v, SEED, fl, 0, lam)
        keep += [t, e]
        views.append(v)
    print("[gpu] largest: mixed meson + meson-meson correlator", flush=True)
    H = G.run_correlator(capi.MESON_MESON, 0, views[0], views[1], src, T, Q, mode=capi.MODE_MIXED)
    H = G.shaped(H, capi.MESON_MESON, 0, src, T, Q)
    del keep
    torch.cuda.empty_cache()
    moms = O.momlist(Q)
    for j in (0, 3, 127):
        t = capi.row_time(src, T, j)
        print(f"[gpu] largest: oracle row j={j} t={t}", flush=True)
        P = [O.synth_prop(L, SEED, fl, 0, lam, t_first=t, nt=1, nthreads=NTH) for fl, lam in enumerate(LAMBDAS)]
        view = (V3, t, 1)
        R = O.correlator(O.MESON_MESON, L, src, T, Q, P[0], P[1], rows=[j], nthreads=NTH, view1=view,
                         view2=view)[:, :, j]
        got = H[:, :, j]
        for name, sl in (("mesons", slice(0, 16)), ("meson-meson", slice(16, 20))):
            e = G.rel_err(got[..., sl], R[..., sl])
            esh = G.shell_rel_err(got[..., sl], R[..., sl], moms)
            pt = G.point_rel_err(got[..., sl], R[..., sl])
            print(f"[gpu] largest j={j} {name}: normwise {e:.2e}, shell-rms {esh:.2e}, pointwise "
                  f"median/p99.9/max {pt[0]:.1e}/{pt[1]:.1e}/{pt[2]:.1e}", flush=True)
            assert e < 1e-6, (j, name, e)
            assert esh < 5e-6, (j, name, esh)
