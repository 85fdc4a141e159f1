"""Contraction tables of the heavy-light tetraquark two-point functions,
derived from the interpolating fields (the reference ships the same tables as
generated data: lib/kernels/PLEGMA_heavy_light_tetraquarks_arrays.cu, made by
python/tetraquarks_v3.py).

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

# tetraquark name -> flavour string of its meson-meson sink field (bbar u bbar d)
TETRAQUARKS = [("bbud", "bt^-up-bt^-dn"), ("bbus", "bt^-up-bt^-sr")]
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


def tables():
    """[(product, [(dataset name, idxs [N, 8], cols [N, 8], vals [N] complex64)])]
    in the reference's product and dataset order"""
    prods = {}
    for name, flavs in TETRAQUARKS:
        ops_out = []
        key = None
        for sink_kind, source_kind, ops in OPERATORS:
            sink = flavs if sink_kind == "meson" else _diquark_form(flavs)
            source = _bar(flavs) if source_kind == "meson" else _bar(_diquark_form(flavs))
            terms = _operator_terms(sink, source)
            maps = []
            for sign, props, fl in terms:
                ms, mc, canon = _default_maps(props, fl)
                maps.append((sign, ms, _colour_slots(props, mc, sink_kind, source_kind)))
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


def write_header(path):
    """C header with the tables: per product the flavours, the operator names,
    counts, and the terms as int16 spins, int16 colour slots and float pairs"""
    lines = ["// generated by plegma_amd/gen_tetra.py from the tetraquark interpolating fields;",
             "// do not edit", "#pragma once", "#include <cstdint>", "namespace plg {"]
    prods = tables()
    lines.append(f"constexpr int TT_NPROD = {len(prods)};")
    lines.append("static const char *const TT_PROD[TT_NPROD] = {" + ", ".join(f'"{p}"' for p, _ in prods) + "};")
    lines.append("static const int TT_NOPS[TT_NPROD] = {" + ", ".join(str(len(o)) for _, o in prods) + "};")
    for pi, (p, ops) in enumerate(prods):
        idx = np.concatenate([o[1] for o in ops]).reshape(-1)
        col = np.concatenate([o[2] for o in ops]).reshape(-1)
        val = np.concatenate([o[3] for o in ops])
        fv = np.stack([val.real, val.imag], -1).astype(np.float32).reshape(-1)
        lines.append(f"static const char *const TT_NAMES_{pi}[] = {{" + ", ".join(f'"{o[0]}"' for o in ops) + "};")
        lines.append(f"static const int16_t TT_COUNT_{pi}[] = {{" + ", ".join(str(len(o[3])) for o in ops) + "};")
        lines.append(f"static const int16_t TT_IDX_{pi}[] = {{" + ",".join(str(int(v)) for v in idx) + "};")
        lines.append(f"static const int16_t TT_COL_{pi}[] = {{" + ",".join(str(int(v)) for v in col) + "};")
        lines.append(f"static const float TT_VAL_{pi}[] = {{" + ",".join(repr(float(v)) + "f" for v in fv) + "};")
    for what in ("NAMES", "COUNT", "IDX", "COL", "VAL"):
        typ = {"NAMES": "const char *const *", "COUNT": "const int16_t *", "IDX": "const int16_t *",
               "COL": "const int16_t *", "VAL": "const float *"}[what]
        lines.append(f"static {typ} const TT_{what}[TT_NPROD] = {{" +
                     ", ".join(f"TT_{what}_{i}" for i in range(len(prods))) + "};")
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
