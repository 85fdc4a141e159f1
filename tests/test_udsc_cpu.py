"""udsc baryon tables and oracle (CPU).

- The term tables generated from the interpolating fields
  (plegma_amd/gen_udsc.py) against the reference's own generated tables,
  through tests/golden/udsc_golden.npz (all 200 operators of the 20 flavour
  products evaluated on seeded site propagators by
  tests/golden/make_udsc_golden.py).
- The oracle's restatement of lib/kernels/PLEGMA_baryons_udsc.cu:86-121
  (oracle/plegma_oracle.c: orc_site_udsc) against the same fixture.
- The operators against an independent Wick-theorem closed form
  (tests/closed_forms.py): O = -sum_(P, G1, G2) Tr[P C(G1, G2^T)] / norm, the
  reference's tables carrying an overall sign -1 relative to that form; up
  to the float32 rounding of the tables' coefficients.
- The product selection of contractBaryonsUDSC (lib/PLEGMA_Correlator.cu:157-175)
  through the library's C ABI (no GPU call).
"""
import os

import numpy as np
import pytest

import closed_forms as cf
from oracle import oracle as O
from plegma_amd import gen_udsc

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = np.load(os.path.join(HERE, "golden", "udsc_golden.npz"))


@pytest.fixture(scope="module")
def products():
    return gen_udsc.products()


def _eval(S, letters, t):
    """numpy evaluation of a term table with the full epsilon tensors"""
    ms = sorted(t)
    idx = np.array(ms)
    v = np.array([t[m] for m in ms])
    A = S[letters[0]][:, idx[:, 0], idx[:, 1]]
    B = S[letters[1]][:, idx[:, 2], idx[:, 3]]
    C3 = S[letters[2]][:, idx[:, 4], idx[:, 5]]
    return np.einsum("abc,def,snad,snbe,sncf->sn", cf.EPS, cf.EPS, A, B, C3, optimize=True) @ v


def test_tables_match_reference(products):
    S = {f: GOLD["props"][i] for i, f in enumerate("udsc")}
    names = list(GOLD["names"])
    assert [p for p, _ in products] == list(GOLD["products"])
    k = 0
    for p, ops in products:
        for n, t in ops:
            assert names[k] == p + ":" + n
            ref = GOLD["values"][k]
            got = _eval(S, p, t)
            assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
            k += 1
    assert k == len(names) == 200


def test_oracle_site_matches_reference(products):
    S = {f: GOLD["props"][i] for i, f in enumerate("udsc")}
    k = 0
    for p, ops in products:
        for site in range(2):
            got = O.udsc_site(S[p[0]][site], S[p[1]][site], S[p[2]][site], ops)
            ref = GOLD["values"][k:k + len(ops), site]
            assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
        k += len(ops)


@pytest.mark.parametrize("baryon", ["proton", "Lambda", "Sigma0", "Xi'c+", "Lambdac+", "Xi*c+", "Delta++",
                                    "Sigma*c++", "Omega-"])
def test_closed_form(products, baryon):
    rng = np.random.default_rng(5)
    S = {f: rng.standard_normal((4, 4, 3, 3)) + 1j * rng.standard_normal((4, 4, 3, 3)) for f in "udsc"}
    full = {"up": "u", "dn": "d", "sr": "s", "ch": "c"}
    name, flavs, coeffs, spin2 = next(b for b in gen_udsc.BARYONS if b[0] == baryon)
    ops = gen_udsc.OPS_HALF if spin2 == 1 else gen_udsc.OPS_3HALF
    norm = {3: 0.5, 2: 1.0, 1: 1.0 / 3.0}[len(coeffs)] if spin2 == 3 else float(sum(c * c for c in coeffs))
    terms = [(c, "".join(full[f] for f in fl.split("-"))) for fl, c in zip(flavs, coeffs)]
    letters = None
    for p, tabs in products:
        for n, t in tabs:
            if n.startswith(gen_udsc.dataset_name(name, "x")[:-1]):
                letters = p
    tabs = dict(dict(products)[letters])
    props = {f: S[f] for f in "udsc"}
    for op, triples in ops.items():
        t = tabs[gen_udsc.dataset_name(name, op)]
        got = _eval({f: S[f][None] for f in "udsc"}, letters, t)[0]
        ref = 0
        for P, G1, G2 in triples:
            Cw = cf.baryon_wick(props, terms, terms, G1, Ab=G2.T)
            ref = ref - np.einsum("mM,mM->", P, Cw) / norm
        assert abs(got - ref) <= 1e-6 * max(abs(ref), 1e-3 * np.abs(S["u"]).max() ** 3), (op, got, ref)


def test_selection():
    """products run for absent flavours and the only_st / only_ch switches"""
    from plegma_amd import capi
    prods = capi.udsc_products()
    dummy = capi.plg_prop()
    for absent in ("", "c", "s", "sc", "u"):
        for only_st in (0, 1):
            for only_ch in (0, 1):
                views = [None if f in absent else dummy for f in "udsc"]
                got, nd = capi.udsc_select(*views, only_st=only_st, only_ch=only_ch)
                want = [i for i, (fl, _) in enumerate(prods)
                        if not any(f in absent for f in fl) and (not only_st or "s" in fl)
                        and (not only_ch or "c" in fl)]
                assert got == want
                assert nd == sum(len(prods[i][1]) for i in want)
