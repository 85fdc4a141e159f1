"""GPU parity of the propagator preparation: the in-place kernels
(plg_prop_rotate_physical, plg_prop_apply_boundaries) against the CPU oracle,
and the preparation fused into the contractions' load stage (plg_prop
rotate_sign / bc_t0) against contractions of prepared copies."""
import numpy as np
import pytest

from oracle import oracle as O
from plegma_amd import capi

import gpu_util as G

pytestmark = pytest.mark.gpu

L1 = [8, 8, 8, 16]
V3 = 512


@pytest.fixture(scope="module", autouse=True)
def ctx():
    G.torch_cuda()
    print("\n[gpu] plg_init", L1, flush=True)
    capi.init(L1, 0)
    yield
    capi.finalize()


def _copy(t, v):
    torch = G.torch_cuda()
    c = t.clone()
    w = capi.plg_prop(c.data_ptr(), v.precision, v.t_first, v.nt, 0, v.comp_stride, v.plane_exp)
    return c, w


@pytest.mark.parametrize("sign", [1, -1])
def test_rotate_f64_matches_oracle(sign):
    t, v = G.make_prop(L1, 321, 1, 0, 0.4)
    ref = np.ascontiguousarray(G.host(t).reshape(144, -1, 2))
    O.prop_rotate(ref, L1, sign)
    print(f"[gpu] rotate f64 sign {sign}", flush=True)
    capi.prop_rotate_physical(v, sign)
    got = G.host(t).reshape(144, -1, 2)
    assert np.abs(got - ref).max() <= 1e-15 * np.abs(ref).max()


def test_rotate_f32_matches_oracle():
    t, v = G.make_prop(L1, 322, 1, 0, 0.4)
    (t32, _), v32 = G.cast_f32(t, v)
    ref = np.ascontiguousarray(G.host(t32).reshape(144, -1, 2))
    O.prop_rotate(ref, L1, 1)
    print("[gpu] rotate f32", flush=True)
    capi.prop_rotate_physical(v32, 1)
    got = G.host(t32).reshape(144, -1, 2)
    assert np.abs(got - ref).max() <= 1e-7 * np.abs(ref).max()


@pytest.mark.parametrize("t_first,nt,t0", [(0, 16, 5), (12, 8, 3), (4, 6, 16)])
def test_boundaries_match_oracle(t_first, nt, t0):
    t, v = G.make_prop(L1, 323, 1, 0, 0.4, t_first=t_first, nt=nt)
    ref = np.ascontiguousarray(G.host(t).reshape(144, -1, 2))
    O.prop_boundaries(ref, L1, t0, t_first)
    print(f"[gpu] boundaries t_first {t_first} nt {nt} t0 {t0}", flush=True)
    capi.prop_apply_boundaries(v, t0)
    assert np.array_equal(G.host(t).reshape(144, -1, 2), ref)


# (kind, mask, two propagators)
KINDS = [(capi.MESONS16, 0, False), (capi.MESONS_NEW, 0, True), (capi.BARYONS, capi.G_NUCL, True),
         (capi.TWOPT, capi.G_NUCL | capi.G_DELTAPP_ALL, False), (capi.MESON_MESON, 0, True)]


@pytest.mark.parametrize("kind,mask,two", KINDS)
def test_fused_preparation_f64(kind, mask, two):
    """contractions of views flagged for preparation == contractions of prepared copies"""
    src, totalT, t0 = [1, 2, 3, 4], 16, 4
    t1, v1 = G.make_prop(L1, 401, 1, src[3], 0.5)
    t2, v2 = G.make_prop(L1, 401, 2, src[3], 0.5) if two else (t1, v1)
    # prepared copies: up rotated with +1, down with -1 (as PLEGMA's drivers do)
    c1, w1 = _copy(t1, v1)
    capi.prop_rotate_physical(w1, 1)
    capi.prop_apply_boundaries(w1, t0)
    if two:
        c2, w2 = _copy(t2, v2)
        capi.prop_rotate_physical(w2, -1)
        capi.prop_apply_boundaries(w2, t0)
    else:
        w2 = w1
    f1 = capi.plg_prop(v1.data, v1.precision, v1.t_first, v1.nt, 0, v1.comp_stride, None, 1, t0)
    f2 = capi.plg_prop(v2.data, v2.precision, v2.t_first, v2.nt, 0, v2.comp_stride, None, -1, t0) if two else f1
    print(f"[gpu] fused preparation kind {kind}", flush=True)
    ref = G.run_correlator(kind, mask, w1, w2, src, totalT, 4)
    got = G.run_correlator(kind, mask, f1, f2, src, totalT, 4)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_fused_preparation_mixed():
    src, totalT, t0 = [0, 5, 1, 9], 16, 12
    t1, v1 = G.make_prop(L1, 402, 1, src[3], 0.5)
    (t32, e32), v32 = G.cast_f32(t1, v1, scaled=True)
    c, w = _copy(t32, v32)
    capi.prop_rotate_physical(w, -1)
    capi.prop_apply_boundaries(w, t0)
    f = capi.plg_prop(v32.data, v32.precision, v32.t_first, v32.nt, 0, v32.comp_stride, v32.plane_exp, -1, t0)
    mask = capi.G_NUCL | capi.G_DELTAPP_ALL
    print("[gpu] fused preparation mixed", flush=True)
    ref = G.run_correlator(capi.TWOPT, mask, w, w, src, totalT, 4, mode=capi.MODE_MIXED)
    got = G.run_correlator(capi.TWOPT, mask, f, f, src, totalT, 4, mode=capi.MODE_MIXED)
    assert np.abs(got - ref).max() <= 1e-9 * np.abs(ref).max()
