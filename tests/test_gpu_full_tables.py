"""GPU parity of the table-driven contractions at the full 32^3 x 64 size,
against the CPU oracle.

udsc baryons (contractBaryonsUDSC, lib/PLEGMA_Correlator.cu:136-188): the
20 flavour products of u, d, s, c propagators (200 operators), one source,
the whole time extent as window, |n|^2 <= 9 (123 momenta), through
plg_baryons_udsc in FP64 and mixed precision.  The oracle restates the
reference's term sums (orc_udsc_field) on sampled rows (time slices): every
operator and momentum of the row.

bcud tetraquarks (contractTetraquarksBCUD, lib/PLEGMA_Correlator.cu:252-310):
the 57 operators of "udcb" (63 232 terms, half of them pairing P1 with P4),
in position space through plg_tetraquarks_bcud; the oracle's literal 81-term
colour sums (orc_site_tetra) cost about 0.1 s per site, so sampled sites of
sampled rows are compared, every operator at each.

Tolerances: FP64 1e-12, mixed 1e-6, normwise over momenta (udsc: per
operator and row) or over operators (bcud: per site and row).
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from plegma_amd import capi

import gpu_util as G

pytestmark = pytest.mark.gpu

L = [32, 32, 32, 64]
V3 = L[0] * L[1] * L[2]
SRC = [5, 17, 9, 21]
TOTALT = 64
Q = 9
NTH = min(os.cpu_count() or 1, 16)
ROWS = [0, 11, 40]       # udsc rows j (time slices (21 + j) % 64)
SITE_ROWS = [3, 50]      # bcud rows
NSITES = 24              # bcud sites per row
LAMBDA = {"u": 0.5, "d": 0.6, "s": 0.8, "c": 1.0, "b": 1.2}


@pytest.fixture(scope="module")
def props():
    """FP64 propagators of u, d, s, c, b (seeded, decay from the source
    slice) and their FP32 copies with per-slice exponents"""
    G.torch_cuda()
    print("\n[gpu] plg_init", L, flush=True)
    capi.init(L, 0)
    out = {}
    for i, f in enumerate("udscb"):
        print(f"[gpu] propagator {f} (FP64 + FP32 copy)", flush=True)
        t, v = G.make_prop(L, 4321, 60 + i, SRC[3], LAMBDA[f])
        keep32 = G.cast_f32(t, v, scaled=True)
        out[f] = (t, v, keep32)
    yield out
    out.clear()
    capi.finalize()


def _slice(props, f, t):
    return G.host_slice(props[f][0], L[3], V3, t)


def _udsc_oracle_row(props, j):
    """the oracle's row j of all 20 products: [200][nmom]"""
    from plegma_amd import gen_udsc
    t = O.row_time(L, SRC, TOTALT, j)
    host = {f: _slice(props, f, t) for f in "udsc"}
    rows = []
    for letters, tabs in gen_udsc.products():
        idx, val, first, count = O.udsc_tables(tabs)
        P = [O._prop(host[f], V3, t, 1) for f in letters]
        nops = len(count)
        fld = np.zeros(V3 * nops * 2)
        ip = O.C.POINTER(O.C.c_int)
        rc = O.lib().orc_udsc_field(O._ivec(L), t, O.C.byref(P[0]), O.C.byref(P[1]), O.C.byref(P[2]),
                                    idx.ctypes.data_as(ip), O._ptr(val), first.ctypes.data_as(ip),
                                    count.ctypes.data_as(ip), nops, O._ptr(fld), NTH)
        assert rc == 0
        fld = fld.reshape(V3, nops, 2)
        rows.append(O.dft(L, SRC[:3], Q, fld[..., 0] + 1j * fld[..., 1], NTH).T)
    return np.concatenate(rows)


@pytest.fixture(scope="module")
def udsc_ref(props):
    print("[gpu] udsc oracle rows", ROWS, flush=True)
    return {j: _udsc_oracle_row(props, j) for j in ROWS}


@pytest.mark.parametrize("mode,tol", [(capi.MODE_F64, 1e-12), (capi.MODE_MIXED, 1e-6)])
def test_udsc_full_size(props, udsc_ref, mode, tol):
    torch = G.torch_cuda()
    moms = capi.momlist(Q)
    views = [props[f][1] if mode == capi.MODE_F64 else props[f][2][1] for f in "udsc"]
    n = capi.baryons_udsc_size(*views, SRC, TOTALT, len(moms))
    assert n == 200 * TOTALT * len(moms)
    out = torch.empty((n, 2), dtype=torch.float64, device="cuda")
    print(f"[gpu] udsc 32^3x64, mode {mode}", flush=True)
    capi.baryons_udsc(*views, SRC, TOTALT, moms, mode, capi.MOMENTUM_SPACE, capi.F64, out)
    H = out.view(200, TOTALT, len(moms), 2)
    for j, ref in udsc_ref.items():
        h = H[:, j].to("cpu").numpy()
        got = h[..., 0] + 1j * h[..., 1]
        err = G.rel_err(got, ref, axis=-1)
        print(f"[gpu] udsc row {j}: normwise {err:.2e}", flush=True)
        assert err < tol


def _bcud_site_ref(S, tabs):
    """the oracle's 57 operators at one site: S = [4] complex [4,4,3,3]"""
    (_, ops), = tabs
    idx = np.ascontiguousarray(np.concatenate([o[1] for o in ops]).astype(np.int32).reshape(-1))
    col = np.ascontiguousarray(np.concatenate([o[2] for o in ops]).astype(np.int32).reshape(-1))
    v = np.concatenate([o[3] for o in ops]).astype(np.complex128)
    val = np.ascontiguousarray(np.stack([v.real, v.imag], -1).reshape(-1))
    cnt = np.array([len(o[3]) for o in ops], dtype=np.int32)
    first = np.concatenate([[0], np.cumsum(cnt)[:-1]]).astype(np.int32)
    raw = [np.ascontiguousarray(np.stack([s.real, s.imag], -1).reshape(-1)) for s in S]
    out = np.zeros(2 * len(cnt))
    ip = O.C.POINTER(O.C.c_int)
    O.lib().orc_site_tetra(*[O._ptr(r) for r in raw], idx.ctypes.data_as(ip), col.ctypes.data_as(ip),
                           O._ptr(val), first.ctypes.data_as(ip), cnt.ctypes.data_as(ip), len(cnt), O._ptr(out))
    return out[0::2] + 1j * out[1::2]


@pytest.fixture(scope="module")
def bcud_ref(props):
    """{row j: (sites, [57][len(sites)])} over the flavours u (LT), c (CH), b (BT)"""
    from plegma_amd import gen_tetra
    tabs = gen_tetra.bcud_tables()
    rng = np.random.default_rng(17)
    out = {}
    for j in SITE_ROWS:
        t = O.row_time(L, SRC, TOTALT, j)
        site = {f: O.prop_to_sites(_slice(props, f, t), L, t, comp_stride=V3, t_first=t) for f in "ucb"}
        xs = np.sort(rng.choice(V3, NSITES, replace=False))
        print(f"[gpu] bcud oracle row {j}: {NSITES} sites", flush=True)
        ref = np.stack([_bcud_site_ref([site["u"][x], site["u"][x], site["c"][x], site["b"][x]], tabs)
                        for x in xs], -1)
        out[j] = (xs, ref)
    return out


@pytest.mark.parametrize("mode,tol", [(capi.MODE_F64, 1e-12), (capi.MODE_MIXED, 1e-6)])
def test_bcud_full_size(props, bcud_ref, mode, tol):
    torch = G.torch_cuda()
    pick = (lambda f: props[f][1]) if mode == capi.MODE_F64 else (lambda f: props[f][2][1])
    views = [pick("u"), None, pick("c"), pick("b")]
    n = capi.tetraquarks_size(*views, SRC, TOTALT, 0, space=capi.POSITION_SPACE, family="bcud")
    assert n == 57 * TOTALT * V3
    out = torch.empty((n, 2), dtype=torch.float64, device="cuda")
    print(f"[gpu] bcud 32^3x64 position space, mode {mode}", flush=True)
    capi.tetraquarks(*views, SRC, TOTALT, [], mode, capi.POSITION_SPACE, capi.F64, out, family="bcud")
    H = out.view(57, TOTALT, V3, 2)
    for j, (xs, ref) in bcud_ref.items():
        idx = torch.tensor(xs, dtype=torch.long, device="cuda")
        h = H[:, j].index_select(1, idx).to("cpu").numpy()
        got = h[..., 0] + 1j * h[..., 1]
        err = G.rel_err(got, ref, axis=0)
        print(f"[gpu] bcud row {j}: normwise over operators {err:.2e}", flush=True)
        assert err < tol
