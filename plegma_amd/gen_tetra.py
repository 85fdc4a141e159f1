"""Contraction tables of the tetraquark two-point functions, derived from the
interpolating fields (the reference ships the same tables as generated data):
- heavy-light bbud / bbus: lib/kernels/PLEGMA_heavy_light_tetraquarks_arrays.cu,
  made by python/tetraquarks_v3.py;
- bcud (J = 1 and J = 0, flavour content (bbar u cbar d) - (bbar d cbar u)):
  lib/kernels/PLEGMA_bcud_tetraquarks_arrays.cu, made by
  python/tetraquarks_v4.py (operator list :105-545, flavour contents
  :549-558, table assembly :1100-1235).

A tetraquark (bbar G1 q1)(bbar G2 q2) -- "meson" form -- or its
diquark-antidiquark regrouping (bbar bbar)(q1 q2) -- "diquark" form -- at the
sink and at the source gives, after Wick contraction, a sum of terms
    v_t * sum_{a0..a3} P1[m0][m1][a_c0][a_c1] P2[m2][m3][a_c2][a_c3]
                     * conj(P3[(m5+2)%4][(m4+2)%4][a_c5][a_c4])
                     * conj(P4[(m7+2)%4][(m6+2)%4][a_c7][a_c6])
(lib/kernels/PLEGMA_heavy_light_tetraquarks.cu:5-61): P1..P4 the propagators of
the flavours sorted as up, dn, sr, ch, bt (the "udbb" product), m0..m7 spin
indices, c0..c7 the colour sum each colour slot runs over (the two quarks of a
meson, or the slots of a diquark, share one sum).  The table of an operator is
its list of distinct (spins, colour slots) keys with the coefficient sums that
do not vanish, in the order the keys first appear: gamma tuples in order,
then Wick terms, then the non-zero entries of the four gammas.

Operator list and gamma tuples: python/tetraquarks_v3.py:31-446; flavour
contents: :451-462; Wick rule and canonical ("default") view: :519-628;
colour slots: :795-842; table assembly: :1040-1120.  Coefficients are stored as
float32, as the reference's float2 arrays.

A field with several flavour orderings is their sum with the given
coefficients; <J Jbar> then sums the Wick terms of every (sink ordering,
source ordering) pair, sink-major, each with the product of the two
coefficients (the order the bcud generator visits them).

Run as a script to write the C header that the library compiles in:
    python -m plegma_amd.gen_tetra plegma_amd/csrc/plg_tetra_tables.h
"""
from __future__ import annotations

import sys

import numpy as np

I = 1j
# gamma basis (lib/kernels/PLEGMA_gammas.cuh:6-41)
GX = np.array([[0, 0, 0, I], [0, 0, I, 0], [0, -I, 0, 0], [-I, 0, 0, 0]])
GY = np.array([[0, 0, 0, 1], [0, 0, -1, 0], [0, -1, 0, 0], [1, 0, 0, 0]], dtype=complex)
GZ = np.array([[0, 0, I, 0], [0, 0, 0, -I], [-I, 0, 0, 0], [0, I, 0, 0]])
GT = np.diag([1, 1, -1, -1]).astype(complex)
G5 = GX @ GY @ GZ @ GT
# the tetraquark generator's charge conjugation is gy gt (without the factor i
# of the baryon tables)
C = GY @ GT
CG5 = C @ G5
GI = {"x": GX, "y": GY, "z": GZ}
GIC = {k: g @ C for k, g in GI.items()}                       # gi C
BARC = {k: GT @ g.conj().T @ GT for k, g in GIC.items()}      # g4 (gi C)^dag g4


def _d(g):
    return g.conj().T


def _meson_meson():
    x, y, z = GX, GY, GZ
    ops = {}
    ops["g5-gj-g5-gj_ll"] = [(G5, g, _d(G5), _d(g)) for g in (x, y, z)]
    ops["g5-gj-gj-g5_ll"] = [(G5, g, _d(g), _d(G5)) for g in (x, y, z)]
    ops["gj-g5-g5-gj_ll"] = [(g, G5, _d(G5), _d(g)) for g in (x, y, z)]
    ops["gj-g5-gj-g5_ll"] = [(g, G5, _d(g), _d(G5)) for g in (x, y, z)]
    # (i, j, k) cyclic with +g5, anticyclic with -g5
    cyc = [("x", "y", "z", 1), ("x", "z", "y", -1), ("y", "z", "x", 1), ("y", "x", "z", -1),
           ("z", "x", "y", 1), ("z", "y", "x", -1)]
    ops["gi-gj-g5-gk_ll"] = [(GI[i], GI[j], _d(s * G5), _d(GI[k])) for i, j, k, s in cyc]
    ops["gi-gj-gk-g5_ll"] = [(GI[i], GI[j], _d(GI[k]), _d(s * G5)) for i, j, k, s in cyc]
    lead = [("z", "x", "y", 1), ("y", "x", "z", -1), ("x", "y", "z", 1), ("z", "y", "x", -1),
            ("y", "z", "x", 1), ("x", "z", "y", -1)]
    ops["g5-gk-gi-gj_ll"] = [(s * G5, GI[k], _d(GI[i]), _d(GI[j])) for k, i, j, s in lead]
    ops["gk-g5-gi-gj_ll"] = [(GI[k], s * G5, _d(GI[i]), _d(GI[j])) for k, i, j, s in lead]
    quad = []
    for i, j in (("x", "y"), ("x", "z"), ("y", "x"), ("y", "z"), ("z", "x"), ("z", "y")):
        quad.append((GI[i], GI[j], _d(GI[i]), _d(GI[j])))
        quad.append((-GI[i], GI[j], _d(GI[j]), _d(GI[i])))
    ops["gi-gj-gk-gl_ll"] = quad
    return ops


def _meson_diquark():
    ops = {}
    ops["g5-gj-gjC-Cg5_ll"] = [(G5, GI[k], BARC[k], -CG5) for k in "xyz"]
    ops["gj-g5-gjC-Cg5_ll"] = [(GI[k], G5, BARC[k], -CG5) for k in "xyz"]
    tri = [("x", "y", "z", 1), ("x", "z", "y", -1), ("y", "z", "x", 1), ("y", "x", "z", -1),
           ("z", "x", "y", 1), ("z", "y", "x", -1)]
    ops["gi-gj-gkC-Cg5_ll"] = [(GI[i], s * GI[j], BARC[k], -CG5) for i, j, k, s in tri]
    return ops


def _diquark_meson():
    ops = {}
    ops["gjC-Cg5-g5-gj_ll"] = [(GIC[k], CG5, _d(G5), _d(GI[k])) for k in "xyz"]
    ops["gjC-Cg5-gj-g5_ll"] = [(GIC[k], CG5, _d(GI[k]), _d(G5)) for k in "xyz"]
    tri = [("z", "x", "y", 1), ("y", "x", "z", -1), ("x", "y", "z", 1), ("z", "y", "x", -1),
           ("y", "z", "x", 1), ("x", "z", "y", -1)]
    ops["gkC-Cg5-gi-gj_ll"] = [(GIC[k], CG5, _d(s * GI[i]), _d(GI[j])) for k, i, j, s in tri]
    return ops


def _diquark_diquark():
    return {"gjC-Cg5-gjC-Cg5_ll": [(GIC[k], CG5, BARC[k], -CG5) for k in "xyz"]}


OPERATORS = [("meson", "meson", _meson_meson()), ("meson", "diquark", _meson_diquark()),
             ("diquark", "meson", _diquark_meson()), ("diquark", "diquark", _diquark_diquark())]


# ---------------------------------------------------------------- bcud operators
# The bcud generator's spin-1 set: every spin-1 form summed over j = x, y, z
# (no suffix), over the (x, y) pairing (suffix _12) and for z alone (suffix _z);
# an odd permutation of the epsilon tensor enters as a minus sign on one gamma.
def _bcud_j1():
    x, y, z = GX, GY, GZ
    d = _d
    mm, md, dm, dd = {}, {}, {}, {}
    for a, b in (("g5", "gj"), ("gj", "g5")):
        for c, e in (("g5", "gj"), ("gj", "g5")):
            def tup(j, k, a=a, b=b, c=c, e=e):
                G1 = G5 if a == "g5" else j
                G2 = G5 if b == "g5" else j
                G3 = G5 if c == "g5" else k
                G4 = G5 if e == "g5" else k
                return (G1, G2, d(G3), d(G4))
            base = f"{a}-{b}-{c}-{e}_ll"
            mm[base] = [tup(g, g) for g in (x, y, z)]
            mm[base + "_12"] = [tup(x, y), tup(y, x)]
            mm[base.replace("gj", "g3") + "_z"] = [tup(z, z)]
    m5 = -G5
    mm["gi-gj-g5-gk_ll"] = [(y, z, d(G5), d(x)), (z, y, d(m5), d(x)), (x, z, d(m5), d(y)),
                            (z, x, d(G5), d(y)), (x, y, d(G5), d(z)), (y, x, d(m5), d(z))]
    mm["gi-gj-g5-gk_ll_12"] = [(y, z, d(G5), d(y)), (z, y, d(m5), d(y)), (x, z, d(m5), d(x)),
                               (z, x, d(G5), d(x))]
    mm["gi-gj-g5-gk_ll_z"] = [(x, y, d(G5), d(z)), (y, x, d(m5), d(z))]
    mm["gi-gj-gk-g5_ll"] = [(y, z, d(x), d(G5)), (z, y, d(x), d(m5)), (x, z, d(y), d(m5)),
                            (z, x, d(y), d(G5)), (x, y, d(z), d(G5)), (y, x, d(z), d(m5))]
    mm["gi-gj-gk-g5_ll_12"] = [(y, z, d(y), d(G5)), (z, y, d(y), d(m5)), (x, z, d(x), d(m5)),
                               (z, x, d(x), d(G5))]
    mm["gi-gj-gk-g5_ll_z"] = [(x, y, d(z), d(G5)), (y, x, d(z), d(m5))]
    mm["g5-gi-gj-gk_ll"] = [(G5, x, d(y), d(z)), (m5, x, d(z), d(y)), (m5, y, d(x), d(z)),
                            (G5, y, d(z), d(x)), (G5, z, d(x), d(y)), (m5, z, d(y), d(x))]
    mm["g5-gi-gj-gk_ll_12"] = [(m5, x, d(x), d(z)), (G5, x, d(z), d(x)), (G5, y, d(y), d(z)),
                               (m5, y, d(z), d(y))]
    mm["g5-gi-gj-gk_ll_z"] = [(G5, z, d(x), d(y)), (m5, z, d(y), d(x))]
    mm["gi-g5-gj-gk_ll"] = [(x, G5, d(y), d(z)), (x, m5, d(z), d(y)), (y, m5, d(x), d(z)),
                            (y, G5, d(z), d(x)), (z, G5, d(x), d(y)), (z, m5, d(y), d(x))]
    mm["gi-g5-gj-gk_ll_12"] = [(x, m5, d(x), d(z)), (x, G5, d(z), d(x)), (y, G5, d(y), d(z)),
                               (y, m5, d(z), d(y))]
    mm["gi-g5-gj-gk_ll_z"] = [(z, G5, d(x), d(y)), (z, m5, d(y), d(x))]
    mm["gi-gj-gi-gj_ll"] = [(y, z, d(y), d(z)), (-y, z, d(z), d(y)), (z, y, d(z), d(y)), (-z, y, d(y), d(z)),
                            (x, z, d(x), d(z)), (-x, z, d(z), d(x)), (z, x, d(z), d(x)), (-z, x, d(x), d(z)),
                            (x, y, d(x), d(y)), (-x, y, d(y), d(x)), (y, x, d(y), d(x)), (-y, x, d(x), d(y))]
    mm["gi-gj-gi-gj_ll_12"] = [(y, z, d(z), d(x)), (-y, z, d(x), d(z)), (-z, y, d(z), d(x)), (z, y, d(x), d(z)),
                               (z, x, d(y), d(z)), (-z, x, d(z), d(y)), (-x, z, d(y), d(z)), (x, z, d(z), d(y))]
    mm["gi-gj-gi-gj_ll_z"] = [(x, y, d(x), d(y)), (-x, y, d(y), d(x)), (y, x, d(y), d(x)), (-y, x, d(x), d(y))]
    # g4 (g_k C)^dag g4 and -C g5 close the diquark of the source
    for a, b in (("g5", "gj"), ("gj", "g5")):
        def tup(j, k, a=a, b=b):
            return (G5 if a == "g5" else GI[j], G5 if b == "g5" else GI[j], BARC[k], -CG5)
        base = f"{a}-{b}-gjC-Cg5_ll"
        md[base] = [tup(k, k) for k in "xyz"]
        md[base + "_12"] = [tup("x", "y"), tup("y", "x")]
        md[base.replace("gj", "g3") + "_z"] = [tup("z", "z")]
    md["gi-gj-gkC-Cg5_ll"] = [(y, z, BARC["x"], -CG5), (z, -y, BARC["x"], -CG5), (x, -z, BARC["y"], -CG5),
                              (z, x, BARC["y"], -CG5), (x, y, BARC["z"], -CG5), (y, -x, BARC["z"], -CG5)]
    md["gi-gj-gkC-Cg5_ll_12"] = [(y, z, BARC["y"], -CG5), (z, -y, BARC["y"], -CG5), (x, -z, BARC["x"], -CG5),
                                 (z, x, BARC["x"], -CG5)]
    md["gi-gj-gkC-Cg5_ll_z"] = [(x, y, BARC["z"], -CG5), (y, -x, BARC["z"], -CG5)]
    for c, e in (("g5", "gj"), ("gj", "g5")):
        def tup(j, k, c=c, e=e):
            return (GIC[j], CG5, d(G5 if c == "g5" else GI[k]), d(G5 if e == "g5" else GI[k]))
        base = f"gjC-Cg5-{c}-{e}_ll"
        dm[base] = [tup(k, k) for k in "xyz"]
        dm[base + "_12"] = [tup("x", "y"), tup("y", "x")]
        dm[base.replace("gj", "g3") + "_z"] = [tup("z", "z")]
    dm["giC-Cg5-gj-gk_ll"] = [(GIC["x"], CG5, d(y), d(z)), (GIC["x"], CG5, d(-z), d(y)),
                              (GIC["y"], CG5, d(-x), d(z)), (GIC["y"], CG5, d(z), d(x)),
                              (GIC["z"], CG5, d(x), d(y)), (GIC["z"], CG5, d(-y), d(x))]
    dm["giC-Cg5-gj-gk_ll_12"] = [(GIC["x"], CG5, d(z), d(x)), (GIC["x"], CG5, d(-x), d(z)),
                                 (GIC["y"], CG5, d(y), d(z)), (GIC["y"], CG5, d(-z), d(y))]
    dm["giC-Cg5-gj-gk_ll_z"] = [(GIC["z"], CG5, d(x), d(y)), (GIC["z"], CG5, d(-y), d(x))]
    dd["gjC-Cg5-gjC-Cg5_ll"] = [(GIC[k], CG5, BARC[k], -CG5) for k in "xyz"]
    dd["gjC-Cg5-gjC-Cg5_ll_12"] = [(GIC["x"], CG5, BARC["y"], -CG5), (GIC["y"], CG5, BARC["x"], -CG5)]
    dd["g3C-Cg5-g3C-Cg5_ll_z"] = [(GIC["z"], CG5, BARC["z"], -CG5)]
    return [("meson", "meson", mm), ("meson", "diquark", md), ("diquark", "meson", dm), ("diquark", "diquark", dd)]


def _bcud_j0():
    """the bcud generator's spin-0 set"""
    d = _d
    G5C = G5 @ C
    BAR5 = GT @ G5C.conj().T @ GT       # g4 (g5 C)^dag g4
    gs = (GX, GY, GZ)
    mm = {"g5-g5-g5-g5_ll": [(G5, G5, d(G5), d(G5))],
          "g5-g5-gj-gj_ll": [(G5, G5, d(g), d(g)) for g in gs],
          "gj-gj-g5-g5_ll": [(g, g, d(G5), d(G5)) for g in gs],
          "gj-gj-gk-gk_ll": [(j, j, d(k), d(k)) for k in gs for j in gs]}
    md = {"g5-g5-g5C-Cg5_ll": [(G5, G5, BAR5, -CG5)],
          "gj-gj-g5C-Cg5_ll": [(g, g, BAR5, -CG5) for g in gs]}
    dm = {"g5C-Cg5-g5-g5_ll": [(G5C, CG5, d(G5), d(G5))],
          "g5C-Cg5-gj-gj_ll": [(G5C, CG5, d(g), d(g)) for g in gs]}
    dd = {"g5C-Cg5-g5C-Cg5_ll": [(G5C, CG5, BAR5, -CG5)]}
    return [("meson", "meson", mm), ("meson", "diquark", md), ("diquark", "meson", dm), ("diquark", "diquark", dd)]


# tetraquark name, flavour orderings of its meson-meson sink field and their
# coefficients, operator sets
TETRAQUARKS = [("bbud", ["bt^-up-bt^-dn"], [1], OPERATORS), ("bbus", ["bt^-up-bt^-sr"], [1], OPERATORS)]
BCUD = [("bcud_J1", ["bt^-up-ch^-dn", "bt^-dn-ch^-up"], [1, -1], _bcud_j1()),
        ("bcud_J0", ["bt^-up-ch^-dn", "bt^-dn-ch^-up"], [1, -1], _bcud_j0())]
ORDER = {"up": 0, "dn": 1, "sr": 2, "ch": 3, "bt": 4}


# ---------------------------------------------------------------- fields
def _parse(flavs):
    """[(flavour, is_antiquark)] of a field string ("bt^" = antiquark)"""
    return [(f[:-1], True) if f.endswith("^") else (f, False) for f in flavs.split("-")]


def _fmt(q):
    return "-".join(f + ("^" if a else "") for f, a in q)


def _bar(flavs):
    """conjugate field: quarks (1, 0, 3, 2) with quark <-> antiquark"""
    q = _parse(flavs)
    return _fmt([(f, not a) for f, a in (q[1], q[0], q[3], q[2])])


def _diquark_form(flavs):
    """(bbar q1)(bbar q2) regrouped as (bbar bbar)(q1 q2): quarks (0, 2, 1, 3)"""
    q = _parse(flavs)
    return _fmt([q[0], q[2], q[1], q[3]])


def _wick(quarks):
    """all pairings of quarks (index, flavour, anti) into propagators
    (quark index, antiquark index), with the fermion sign of each"""
    q0, rest = quarks[0], quarks[1:]
    out = []
    for i, q in enumerate(rest):
        if q[1] != q0[1] or q[2] == q0[2]:
            continue
        prop = (q0[0], q[0]) if q[2] else (q[0], q0[0])
        sign = (1 if i % 2 == 0 else -1) * (1 if q[2] else -1)
        others = rest[:i] + rest[i + 1:]
        if not others:
            out.append(([prop], sign))
        else:
            for props, s in _wick(others):
                out.append(([prop] + props, sign * s))
    return out


def _operator_terms(sink, source):
    """Wick terms of <J Jbar>: (sign, propagators, flavour of each)"""
    quarks = [(i, f, a) for i, (f, a) in enumerate(_parse(sink) + _parse(source))]
    flav = {i: f for i, f, _ in quarks}
    return [(s, props, [flav[p[0]] for p in props]) for props, s in _wick(quarks)]


def _default_maps(props, flavs):
    """index maps of one Wick term onto the canonical view: the propagators
    sorted by flavour, propagator k with spins (2k, 2k+1) and colours
    (k, 4 + k); returns ({original spin: canonical}, {original colour: canonical})"""
    canon = sorted(flavs, key=lambda f: ORDER[f])
    free = list(range(len(canon)))
    ms, mc = {}, {}
    for (qi, ai), f in zip(props, flavs):
        k = next(j for j in free if canon[j] == f)
        free.remove(k)
        ms[qi], ms[ai] = 2 * k, 2 * k + 1
        mc[qi], mc[ai] = k, 4 + k
    return ms, mc, canon


# canonical colour index -> colour slot of a table term: slots (2k, 2k+1) are
# the (quark, antiquark) colours of propagator k
_SLOT = {0: 0, 4: 1, 1: 2, 5: 3, 2: 4, 6: 5, 3: 6, 7: 7}


def _colour_slots(props, mc, sink_kind, source_kind):
    """the colour sum (0..3) each slot runs over: the quarks of one meson
    share a sum, and so do quarks 0, 2 and quarks 1, 3 of a diquark pair"""
    slots = [0] * 8
    for p in props:
        for q in p:
            if q < 4:
                c = q // 2 if sink_kind == "meson" else q % 2
            else:
                c = q // 2 if source_kind == "meson" else q % 2 + 2
            slots[_SLOT[mc[q]]] = c
    return tuple(slots)


def _nonzero(mats):
    """non-zero entries of the outer product of four 4x4 matrices, the first
    matrix outermost: ([r0, c0, r1, c1, r2, c2, r3, c3], value)"""
    out = [([], 1)]
    for m in mats:
        nz = list(zip(*np.nonzero(m)))
        out = [(ix + [int(r), int(c)], v * m[r, c]) for ix, v in out for r, c in nz]
    return out


def _tables(tetraquarks):
    prods = {}
    for name, flavs_list, coeffs, opsets in tetraquarks:
        ops_out = []
        key = None
        for sink_kind, source_kind, ops in opsets:
            maps = []
            for fs, cs in zip(flavs_list, coeffs):
                sink = fs if sink_kind == "meson" else _diquark_form(fs)
                for fb, cb in zip(flavs_list, coeffs):
                    source = _bar(fb) if source_kind == "meson" else _bar(_diquark_form(fb))
                    for sign, props, fl in _operator_terms(sink, source):
                        ms, mc, canon = _default_maps(props, fl)
                        maps.append((sign * cs * cb, ms, _colour_slots(props, mc, sink_kind, source_kind)))
                        key = "".join(f[0] for f in canon)
            for opname, gammas in ops.items():
                acc = {}
                for g in gammas:
                    nz = _nonzero(g)
                    for sign, ms, slots in maps:
                        for ix, v in nz:
                            spins = [0] * 8
                            for orig, canon_pos in ms.items():
                                spins[canon_pos] = ix[orig]
                            k = (tuple(spins), slots)
                            acc[k] = acc.get(k, 0) + v * sign
                acc = {k: v for k, v in acc.items() if not np.isclose(v, 0)}
                idxs = np.array([k[0] for k in acc], dtype=np.int16).reshape(-1, 8)
                cols = np.array([k[1] for k in acc], dtype=np.int16).reshape(-1, 8)
                vals = np.array(list(acc.values()), dtype=np.complex128).astype(np.complex64)
                ops_out.append((name + "/" + opname.replace("-", "_"), idxs, cols, vals))
        prods.setdefault(key, []).extend(ops_out)
    return list(prods.items())


def tables():
    """heavy-light tables: [(product, [(dataset name, idxs [N, 8], cols [N, 8],
    vals [N] complex64)])] in the reference's product and dataset order"""
    return _tables(TETRAQUARKS)


def bcud_tables():
    """bcud tables (product "udcb"), same layout"""
    return _tables(BCUD)


def _emit(lines, tag, prods):
    """the tables of one set as C arrays with names TT_* (heavy-light) or
    TB_* (bcud)"""
    lines.append(f"constexpr int {tag}_NPROD = {len(prods)};")
    lines.append(f"static const char *const {tag}_PROD[{tag}_NPROD] = {{" + ", ".join(f'"{p}"' for p, _ in prods) + "};")
    lines.append(f"static const int {tag}_NOPS[{tag}_NPROD] = {{" + ", ".join(str(len(o)) for _, o in prods) + "};")
    for pi, (p, ops) in enumerate(prods):
        idx = np.concatenate([o[1] for o in ops]).reshape(-1)
        col = np.concatenate([o[2] for o in ops]).reshape(-1)
        val = np.concatenate([o[3] for o in ops])
        fv = np.stack([val.real, val.imag], -1).astype(np.float32).reshape(-1)
        lines.append(f"static const char *const {tag}_NAMES_{pi}[] = {{" + ", ".join(f'"{o[0]}"' for o in ops) + "};")
        lines.append(f"static const int16_t {tag}_COUNT_{pi}[] = {{" + ", ".join(str(len(o[3])) for o in ops) + "};")
        lines.append(f"static const int16_t {tag}_IDX_{pi}[] = {{" + ",".join(str(int(v)) for v in idx) + "};")
        lines.append(f"static const int16_t {tag}_COL_{pi}[] = {{" + ",".join(str(int(v)) for v in col) + "};")
        lines.append(f"static const float {tag}_VAL_{pi}[] = {{" + ",".join(repr(float(v)) + "f" for v in fv) + "};")
    for what in ("NAMES", "COUNT", "IDX", "COL", "VAL"):
        typ = {"NAMES": "const char *const *", "COUNT": "const int16_t *", "IDX": "const int16_t *",
               "COL": "const int16_t *", "VAL": "const float *"}[what]
        lines.append(f"static {typ} const {tag}_{what}[{tag}_NPROD] = {{" +
                     ", ".join(f"{tag}_{what}_{i}" for i in range(len(prods))) + "};")


def write_header(path):
    """C header with the tables: per product the flavours, the operator names,
    counts, and the terms as int16 spins, int16 colour slots and float pairs;
    TT_* the heavy-light set, TB_* the bcud set"""
    lines = ["// generated by plegma_amd/gen_tetra.py from the tetraquark interpolating fields;",
             "// do not edit", "#pragma once", "#include <cstdint>", "namespace plg {"]
    _emit(lines, "TT", tables())
    _emit(lines, "TB", bcud_tables())
    lines.append("}  // namespace plg")
    text = "\n".join(lines) + "\n"
    try:
        with open(path) as f:
            if f.read() == text:
                return
    except OSError:
        pass
    with open(path, "w") as f:
        f.write(text)


if __name__ == "__main__":
    write_header(sys.argv[1])
