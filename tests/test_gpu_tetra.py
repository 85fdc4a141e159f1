"""Four-propagator table contractions on the GPU (plg_tetraquark_table, the
counterpart of contract_props in lib/kernels/PLEGMA_heavy_light_tetraquarks.cu:6-62
with the reference's table format) against the oracle's literal 81-term
colour sums, on random tables in the reference's four colour patterns."""
import numpy as np
import pytest

from oracle import oracle as O
from plegma_amd import capi

import gpu_util as G

pytestmark = pytest.mark.gpu

L = [4, 4, 4, 8]
SRC = [2, 1, 3, 6]
TOTALT = 5
Q2 = 2
PATTERNS = [(0, 2, 1, 3, 2, 0, 3, 1), (0, 3, 1, 2, 3, 0, 2, 1), (0, 2, 1, 3, 3, 0, 2, 1), (0, 3, 1, 2, 2, 0, 3, 1)]
# the further patterns of the bcud tables: P1-P3 / P2-P4 pairings with other
# colour labels, and the P1-P4 / P2-P3 pairings (a second launch that adds)
SWAPPED = [(0, 2, 1, 3, 3, 1, 2, 0), (0, 3, 1, 2, 2, 1, 3, 0), (0, 3, 1, 2, 3, 1, 2, 0), (0, 2, 1, 3, 2, 1, 3, 0)]
BCUD_PATTERNS = PATTERNS + [(1, 3, 0, 2, 3, 1, 2, 0), (1, 2, 0, 3, 2, 1, 3, 0), (1, 2, 0, 3, 3, 1, 2, 0),
                            (1, 3, 0, 2, 2, 1, 3, 0)] + SWAPPED


@pytest.fixture(scope="module", autouse=True)
def ctx():
    G.torch_cuda()
    print("\n[gpu] plg_init", L, flush=True)
    capi.init(L, 0)
    yield
    capi.finalize()


@pytest.fixture(scope="module")
def props(ctx):
    out = []
    for i, lam in enumerate((0.5, 0.6, 1.1, 1.3)):
        t, v = G.make_prop(L, 91, 30 + i, SRC[3], lam)
        out.append((t, v, G.host(t)))
    return out


def _table(rng, nops=5, nterm=60, patterns=PATTERNS):
    idxs, cols, vals, counts = [], [], [], []
    coefs = [1, -1, 1j, -0.5, 0.25 + 0.5j, 1 / 3]
    for k in range(nops):
        n = int(rng.integers(nterm // 2, nterm))
        counts.append(n)
        for _ in range(n):
            idxs.append(rng.integers(0, 4, size=8))
            cols.append(patterns[rng.integers(len(patterns))])
            v = complex(coefs[rng.integers(len(coefs))])
            vals.append([np.float32(v.real), np.float32(v.imag)])
    return (np.array(idxs, dtype=np.int16), np.array(cols, dtype=np.int16), np.array(vals, dtype=np.float32),
            np.array(counts, dtype=np.int16))


def _run(views, tab, mode=capi.MODE_F64, space=capi.MOMENTUM_SPACE, out_prec=capi.F64):
    torch = G.torch_cuda()
    moms = capi.momlist(Q2) if space == capi.MOMENTUM_SPACE else []
    inner = len(moms) if space == capi.MOMENTUM_SPACE else L[0] * L[1] * L[2]
    idxs, cols, vals, counts = tab
    dt = torch.float64 if out_prec == capi.F64 else torch.float32
    out = torch.full((len(counts) * TOTALT * inner, 2), float("nan"), dtype=dt, device="cuda")
    print(f"[gpu] tetraquark table: {len(counts)} operators, {len(idxs)} terms, mode {mode}, space {space}",
          flush=True)
    capi.tetraquark_table(*views, idxs, cols, vals, counts, SRC, TOTALT, moms, mode, space, out_prec, out)
    torch.cuda.synchronize()
    h = out.to("cpu", torch.float64).numpy()
    return (h[:, 0] + 1j * h[:, 1]).reshape(len(counts), TOTALT, inner)


def _oracle(props, tab, space="momentum", hosts=None, views=None):
    idxs, cols, vals, counts = tab
    v = vals[:, 0].astype(np.float64) + 1j * vals[:, 1].astype(np.float64)
    return O.tetra_correlator(L, SRC, TOTALT, Q2, hosts or [p[2] for p in props], idxs, cols, v, counts,
                              space=space, views=views)


def test_tetra_random_f64(props):
    tab = _table(np.random.default_rng(4))
    ref = _oracle(props, tab)
    got = _run([p[1] for p in props], tab)
    assert G.rel_err(got, ref, axis=-1) < 1e-12


def test_tetra_repeated_views(props):
    """one propagator in two slots (as propLT serves u and d)"""
    tab = _table(np.random.default_rng(5), nops=3)
    P = [props[0], props[0], props[2], props[2]]
    ref = _oracle(P, tab)
    got = _run([p[1] for p in P], tab)
    assert G.rel_err(got, ref, axis=-1) < 1e-12


def test_tetra_mixed(props):
    tab = _table(np.random.default_rng(6), nops=4)
    ref = _oracle(props, tab)
    keep = [G.cast_f32(p[0], p[1], scaled=True) for p in props]
    got = _run([k[1] for k in keep], tab, mode=capi.MODE_MIXED)
    assert G.rel_err(got, ref, axis=-1) < 1e-6


def test_tetra_refuses_other_patterns(props):
    torch = G.torch_cuda()
    out = torch.zeros((TOTALT * 33, 2), dtype=torch.float64, device="cuda")
    idxs = np.zeros((1, 8), dtype=np.int16)
    vals = np.ones((1, 2), dtype=np.float32)
    for cols, code in (((0, 0, 1, 1, 2, 0, 3, 1), "UNSUPPORTED"), ((0, 2, 1, 3, 1, 0, 3, 2), "UNSUPPORTED"),
                       ((0, 2, 1, 3, 2, 0, 3, 5), "INVALID")):
        with pytest.raises(capi.PlgError, match=code):
            capi.tetraquark_table(*[p[1] for p in props], idxs, np.array([cols], dtype=np.int16), vals,
                                  np.ones(1, dtype=np.int16), SRC, TOTALT, capi.momlist(Q2), capi.MODE_F64,
                                  capi.MOMENTUM_SPACE, capi.F64, out)


def test_tetra_position_space_f32_output(props):
    tab = _table(np.random.default_rng(7), nops=3, nterm=30)
    ref = _oracle(props, tab, space="position")
    got = _run([p[1] for p in props], tab, space=capi.POSITION_SPACE, out_prec=capi.F32)
    assert G.rel_err(got, ref, axis=-1) < 1e-6


def test_tetra_window_views(props):
    """propagators stored over the source's window only (t_first, nt): the
    rows of the correlator are t = 6, 7, 0, 1, 2"""
    torch = G.torch_cuda()
    tab = _table(np.random.default_rng(8), nops=3, nterm=30)
    V3 = L[0] * L[1] * L[2]
    nt, tf = 5, 6
    views, hosts, keep = [], [], []
    for i, lam in enumerate((0.5, 0.6, 1.1, 1.3)):
        t = torch.empty((144, nt * V3, 2), dtype=torch.float64, device="cuda")
        v = capi.plg_prop(t.data_ptr(), capi.F64, tf, nt, 0, nt * V3)
        capi.synth_prop(v, 91, 30 + i, SRC[3], lam)
        keep.append(t)
        views.append(v)
        hosts.append(G.host(t))
    ref = _oracle(props, tab, hosts=hosts, views=[(nt * V3, tf, nt)] * 4)
    got = _run(views, tab)
    assert G.rel_err(got, ref, axis=-1) < 1e-12


def test_tetra_fused_prep(props):
    """rotation to the physical basis (both signs) and the boundary sign of
    the four views applied on load, against the oracle on prepared copies"""
    tab = _table(np.random.default_rng(9), nops=4, nterm=40)
    hosts, views = [], []
    for (t, v, h), sign in zip(props, (1, -1, 1, -1)):
        w = capi.plg_prop(v.data, v.precision, v.t_first, v.nt, 0, v.comp_stride)
        w.rotate_sign = sign
        w.bc_t0 = SRC[3]
        views.append(w)
        hp = h.copy().reshape(144, -1, 2)
        O.prop_rotate(hp, L, sign)
        O.prop_boundaries(hp, L, SRC[3])
        hosts.append(hp.reshape(-1))
    ref = _oracle(props, tab, hosts=hosts)
    got = _run(views, tab)
    assert G.rel_err(got, ref, axis=-1) < 1e-12


def _builtin_ref(fl, moms_q2=Q2, only_st=False, only_ch=False):
    """the oracle's 81-term sums over the generated tables of every product
    whose flavours are present (None = absent), datasets in order"""
    from plegma_amd import gen_tetra
    names, refs = [], []
    for p, ops in gen_tetra.tables():
        if any(fl[f] is None for f in p) or (only_st and "s" not in p) or (only_ch and "c" not in p):
            continue
        idx = np.concatenate([o[1] for o in ops])
        col = np.concatenate([o[2] for o in ops])
        val = np.concatenate([o[3] for o in ops]).astype(np.complex128)
        refs.append(O.tetra_correlator(L, SRC, TOTALT, moms_q2, [fl[f] for f in p], idx, col, val,
                                       [len(o[3]) for o in ops]))
        names += [o[0] for o in ops]
    return names, (np.concatenate(refs) if refs else np.zeros((0, TOTALT, len(O.momlist(moms_q2)))))


def _run_builtin(lt, st, ch, bt, mode=capi.MODE_F64, only_st=False, only_ch=False):
    torch = G.torch_cuda()
    moms = capi.momlist(Q2)
    n = capi.tetraquarks_size(lt, st, ch, bt, SRC, TOTALT, len(moms), only_st=only_st, only_ch=only_ch)
    out = torch.full((max(n, 1), 2), float("nan"), dtype=torch.float64, device="cuda")
    print(f"[gpu] built-in tetraquarks: {n} elements, mode {mode}", flush=True)
    capi.tetraquarks(lt, st, ch, bt, SRC, TOTALT, moms, mode, capi.MOMENTUM_SPACE, capi.F64, out,
                     only_st=only_st, only_ch=only_ch)
    h = G.host(out).reshape(-1, 2)[:n]
    return (h[:, 0] + 1j * h[:, 1]).reshape(-1, TOTALT, len(moms))


def test_builtin_tetraquarks_products():
    prods = capi.tetra_products()
    from plegma_amd import gen_tetra
    assert [(p, [o[0] for o in ops]) for p, ops in gen_tetra.tables()] == prods
    assert [p for p, _ in prods] == ["udbb", "usbb"]


@pytest.mark.parametrize("mode,tol", [(capi.MODE_F64, 1e-12), (capi.MODE_MIXED, 1e-6)])
def test_builtin_tetraquarks(props, mode, tol):
    """plg_tetraquarks with LT, ST and BT present and CH absent: both
    products ("udbb" with LT twice, "usbb"), 32 datasets, against the oracle
    over the generated tables (which equal the reference's)"""
    lt, st, bt = props[0], props[1], props[3]
    fl = {"u": lt[2], "d": lt[2], "s": st[2], "c": None, "b": bt[2]}
    names, ref = _builtin_ref(fl)
    assert len(names) == 32
    if mode == capi.MODE_F64:
        views = [lt[1], st[1], None, bt[1]]
    else:
        keep = [G.cast_f32(p[0], p[1], scaled=True) for p in (lt, st, bt)]
        views = [keep[0][1], keep[1][1], None, keep[2][1]]
    got = _run_builtin(*views, mode=mode)
    assert got.shape == ref.shape
    assert G.rel_err(got, ref, axis=-1) < tol


def test_builtin_tetraquarks_selection(props):
    """the reference's product selection (lib/PLEGMA_Correlator.cu:209-233):
    ST absent drops "usbb", only_st keeps it alone, BT absent or only_ch
    (no charm product) leaves nothing"""
    lt, st, bt = props[0][1], props[1][1], props[3][1]
    assert capi.tetraquarks_select(lt, None, None, bt) == ([0], 16)
    assert capi.tetraquarks_select(lt, st, None, bt, only_st=True) == ([1], 16)
    assert capi.tetraquarks_select(lt, st, None, None) == ([], 0)
    assert capi.tetraquarks_select(lt, st, None, bt, only_ch=True) == ([], 0)
    fl = {"u": props[0][2], "d": props[0][2], "s": None, "c": None, "b": props[3][2]}
    names, ref = _builtin_ref(fl)
    got = _run_builtin(lt, None, None, bt)
    assert len(names) == 16 and G.rel_err(got, ref, axis=-1) < 1e-12


@pytest.mark.parametrize("patterns", [BCUD_PATTERNS, SWAPPED], ids=["all", "swapped-only"])
def test_tetra_swapped_pairings_f64(props, patterns):
    """tables with terms that pair the colours of P1 with P4 and P2 with P3:
    packed for the views P1, P2, P4, P3 and added by a second launch; with
    swapped terms only the first launch writes zeros"""
    tab = _table(np.random.default_rng(10), nops=6, nterm=50, patterns=patterns)
    ref = _oracle(props, tab)
    got = _run([p[1] for p in props], tab)
    assert G.rel_err(got, ref, axis=-1) < 1e-12


def test_tetra_swapped_pairings_mixed_position(props):
    tab = _table(np.random.default_rng(11), nops=5, nterm=40, patterns=BCUD_PATTERNS)
    keep = [G.cast_f32(p[0], p[1], scaled=True) for p in props]
    ref = _oracle(props, tab)
    got = _run([k[1] for k in keep], tab, mode=capi.MODE_MIXED)
    assert G.rel_err(got, ref, axis=-1) < 1e-6
    ref = _oracle(props, tab, space="position")
    got = _run([p[1] for p in props], tab, space=capi.POSITION_SPACE)
    assert G.rel_err(got, ref, axis=-1) < 1e-12


def _bcud_ref(fl):
    from plegma_amd import gen_tetra
    (p, ops), = gen_tetra.bcud_tables()
    idx = np.concatenate([o[1] for o in ops])
    col = np.concatenate([o[2] for o in ops])
    val = np.concatenate([o[3] for o in ops]).astype(np.complex128)
    ref = O.tetra_correlator(L, SRC, TOTALT, Q2, [fl[f] for f in p], idx, col, val, [len(o[3]) for o in ops])
    return [o[0] for o in ops], ref


def test_builtin_bcud_products():
    from plegma_amd import gen_tetra
    prods = capi.tetra_products("bcud")
    assert [(p, [o[0] for o in ops]) for p, ops in gen_tetra.bcud_tables()] == prods
    assert prods[0][0] == "udcb" and len(prods[0][1]) == 57


@pytest.mark.parametrize("mode,tol", [(capi.MODE_F64, 1e-12), (capi.MODE_MIXED, 1e-6)])
def test_builtin_bcud(props, mode, tol):
    """plg_tetraquarks_bcud with LT, CH and BT (ST absent): the 57 datasets of
    "udcb" against the oracle over the generated tables (which equal the
    reference's)"""
    torch = G.torch_cuda()
    lt, ch, bt = props[0], props[2], props[3]
    names, ref = _bcud_ref({"u": lt[2], "d": lt[2], "c": ch[2], "b": bt[2]})
    if mode == capi.MODE_F64:
        views = [lt[1], None, ch[1], bt[1]]
    else:
        keep = [G.cast_f32(p[0], p[1], scaled=True) for p in (lt, ch, bt)]
        views = [keep[0][1], None, keep[1][1], keep[2][1]]
    moms = capi.momlist(Q2)
    n = capi.tetraquarks_size(*views, SRC, TOTALT, len(moms), family="bcud")
    assert n == ref.size
    out = torch.full((n, 2), float("nan"), dtype=torch.float64, device="cuda")
    print(f"[gpu] bcud tetraquarks: {n} elements, mode {mode}", flush=True)
    capi.tetraquarks(*views, SRC, TOTALT, moms, mode, capi.MOMENTUM_SPACE, capi.F64, out, family="bcud")
    h = G.host(out).reshape(-1, 2)
    got = (h[:, 0] + 1j * h[:, 1]).reshape(ref.shape)
    assert G.rel_err(got, ref, axis=-1) < tol


def test_builtin_bcud_selection(props):
    """the reference's selection (lib/PLEGMA_Correlator.cu:270-293): the
    product needs LT, CH and BT; only_st drops it (no strange quark)"""
    lt, st, ch, bt = (p[1] for p in props)
    assert capi.tetraquarks_select(lt, None, ch, bt, family="bcud") == ([0], 57)
    assert capi.tetraquarks_select(lt, st, None, bt, family="bcud") == ([], 0)
    assert capi.tetraquarks_select(lt, st, ch, bt, only_st=True, family="bcud") == ([], 0)
    assert capi.tetraquarks_select(lt, st, ch, bt, only_ch=True, family="bcud") == ([0], 57)
