"""Contraction tables of the udsc baryon two-point functions, derived from the
interpolating fields (the reference ships the same tables as generated data:
lib/kernels/PLEGMA_baryons_udsc_arrays.cu, made by python/baryons.py).

A baryon interpolator J = sum_k c_k eps_abc (f1_a^T G f2_b) f3_c and its
conjugate Jbar = sum_k c_k eps (f3bar_c) (f1bar_a G' f2bar_b) give, after Wick
contraction, a sum over terms
    v_t * sum_{eps, eps'} sgn sgn' S1[m0][m1][a][a'] S2[m2][m3][b][b'] S3[m4][m5][c][c']
where S1, S2, S3 are the propagators of the flavours in the order up, dn, sr,
ch (python/baryons.py:378-392), m0..m5 spin indices and v_t complex.  Every
Wick term is mapped onto that canonical colour structure with the parity of
its colour permutation (the epsilon tensors are antisymmetric), and onto the
canonical spin slots; the gamma structures contribute their non-zero entries.

Operator sets and baryon list: python/baryons.py:83-287.  Normalisation:
python/baryons.py:484-498.  Coefficients are rounded to float32, as the
reference stores them (float2 arrays) and multiplies FP64 results by them.

Run as a script to write the C header that the library compiles in:
    python -m plegma_amd.gen_udsc plegma_amd/csrc/plg_udsc_tables.h
"""
from __future__ import annotations

import itertools
import sys

import numpy as np

I = 1j
FLAVOURS = ("up", "dn", "sr", "ch")          # canonical order of S1, S2, S3

# gamma basis (lib/kernels/PLEGMA_gammas.cuh:6-41)
GX = np.array([[0, 0, 0, I], [0, 0, I, 0], [0, -I, 0, 0], [-I, 0, 0, 0]])
GY = np.array([[0, 0, 0, 1], [0, 0, -1, 0], [0, -1, 0, 0], [1, 0, 0, 0]], dtype=complex)
GZ = np.array([[0, 0, I, 0], [0, 0, 0, -I], [-I, 0, 0, 0], [0, I, 0, 0]])
GT = np.diag([1, 1, -1, -1]).astype(complex)
ONE = np.eye(4, dtype=complex)
G5 = GX @ GY @ GZ @ GT
C = I * GY @ GT
CG5 = C @ G5
IGTCG5 = I * GT @ CG5
PP, PM = (ONE + GT) / 2, (ONE - GT) / 2
CGI = [C @ GX, C @ GY, C @ GZ]
GI = [GX, GY, GZ]

# (projector, diquark of J, diquark of Jbar) triples summed per operator
OPS_HALF = {
    "Pp-Cg5-Cg5": [(PP, CG5, CG5)],
    "Pm-Cg5-Cg5": [(PM, CG5, CG5)],
    "Pp-igtCg5-igtCg5": [(PP, IGTCG5, IGTCG5)],
    "Pm-igtCg5-igtCg5": [(PM, IGTCG5, IGTCG5)],
    "g5Ppg5-C-C": [(G5 @ PP @ G5, C, C)],
    "g5Pmg5-C-C": [(G5 @ PM @ G5, C, C)],
}
_PAIRS = [(0, 1), (0, 2), (1, 2), (1, 0), (2, 0), (2, 1)]
OPS_3HALF = {
    "Pp-Cgi-Cgi": [(PP, CGI[i], CGI[i]) for i in range(3)],
    "Pm-Cgi-Cgi": [(PM, CGI[i], CGI[i]) for i in range(3)],
    "Ppgigj-Cgi-Cgj": [(PP @ GI[i] @ GI[j], CGI[i], CGI[j]) for i, j in _PAIRS],
    "Pmgigj-Cgi-Cgj": [(PM @ GI[i] @ GI[j], CGI[i], CGI[j]) for i, j in _PAIRS],
}

# name: (flavour strings f1-f2-f3, coefficients, spin x 2)
BARYONS = [
    ("Xicc++", ["ch-up-ch"], [1], 1), ("Xicc+", ["ch-dn-ch"], [1], 1), ("Omegacc+", ["ch-sr-ch"], [1], 1),
    ("Sigmac++", ["up-ch-up"], [1], 1), ("Sigmac+", ["up-ch-dn", "dn-ch-up"], [1, 1], 1),
    ("Sigmac0", ["dn-ch-dn"], [1], 1), ("Xi'c+", ["up-ch-sr", "sr-ch-up"], [1, 1], 1),
    ("Xi'c0", ["dn-ch-sr", "sr-ch-dn"], [1, 1], 1), ("Omegac0", ["sr-ch-sr"], [1], 1),
    ("Lambdac+", ["up-dn-ch", "up-ch-dn", "dn-ch-up"], [2, 1, -1], 1),
    ("Xic+", ["sr-up-ch", "sr-ch-up", "up-ch-sr"], [2, 1, -1], 1),
    ("Xic0", ["sr-dn-ch", "sr-ch-dn", "dn-ch-sr"], [2, 1, -1], 1),
    ("proton", ["up-dn-up"], [1], 1), ("neutron", ["dn-up-dn"], [1], 1),
    ("Lambda", ["up-dn-sr", "up-sr-dn", "dn-sr-up"], [2, 1, -1], 1),
    ("Sigma+", ["up-sr-up"], [1], 1), ("Sigma0", ["up-sr-dn", "dn-sr-up"], [1, 1], 1),
    ("Sigma-", ["dn-sr-dn"], [1], 1), ("Xi0", ["sr-up-sr"], [1], 1), ("Xi-", ["sr-dn-sr"], [1], 1),
    ("Omegaccc++", ["ch-ch-ch"], [1], 3), ("Xi*cc++", ["ch-up-ch", "ch-ch-up"], [2, 1], 3),
    ("Xi*cc+", ["ch-dn-ch", "ch-ch-dn"], [2, 1], 3), ("Omega*cc+", ["ch-sr-ch", "ch-ch-sr"], [2, 1], 3),
    ("Sigma*c++", ["up-up-ch", "ch-up-up"], [1, 2], 3),
    ("Sigma*c+", ["up-dn-ch", "dn-ch-up", "ch-up-dn"], [1, 1, 1], 3),
    ("Sigma*c0", ["dn-dn-ch", "ch-dn-dn"], [1, 2], 3),
    ("Xi*c+", ["up-sr-ch", "sr-ch-up", "ch-up-sr"], [1, 1, 1], 3),
    ("Xi*c0", ["dn-sr-ch", "sr-ch-dn", "ch-dn-sr"], [1, 1, 1], 3),
    ("Omega*c0", ["sr-ch-sr", "sr-sr-ch"], [2, 1], 3),
    ("Delta++", ["up-up-up"], [1], 3), ("Delta+", ["up-dn-up", "up-up-dn"], [2, 1], 3),
    ("Delta0", ["dn-up-dn", "dn-dn-up"], [2, 1], 3), ("Delta-", ["dn-dn-dn"], [1], 3),
    ("Sigma*+", ["up-up-sr", "sr-up-up"], [1, 2], 3),
    ("Sigma*0", ["up-dn-sr", "dn-sr-up", "sr-up-dn"], [1, 1, 1], 3),
    ("Sigma*-", ["dn-dn-sr", "sr-dn-dn"], [1, 2], 3),
    ("Xi*0", ["sr-up-sr", "sr-sr-up"], [2, 1], 3), ("Xi*-", ["sr-dn-sr", "sr-sr-dn"], [2, 1], 3),
    ("Omega-", ["sr-sr-sr"], [1], 3),
]


def _contractions(flav):
    """Wick contractions of the six fields q0 q1 q2 qbar3 qbar4 qbar5 (slot i
    has spin and colour index i).  Each quark is paired with an antiquark of its
    flavour; the sign is that of the permutation bringing every pair adjacent.
    Returns [(sign, [(quark slot, antiquark slot), ...])] with the pairs in the
    order the quarks appear."""
    out = []
    quarks = [0, 1, 2]
    anti = [3, 4, 5]
    for perm in itertools.permutations(anti):
        if any(flav[q] != flav[a] for q, a in zip(quarks, perm)):
            continue
        # q0 qb3 q1 qb4 q2 qb5 -> sign of pairing (q0 perm0)(q1 perm1)(q2 perm2):
        # reorder the sequence 0 1 2 3 4 5 into 0 p0 1 p1 2 p2
        seq = [0, perm[0], 1, perm[1], 2, perm[2]]
        out.append((_parity(seq), list(zip(quarks, perm))))
    return out


def _parity(seq):
    s = 1
    seq = list(seq)
    for i in range(len(seq)):
        while seq[i] != i:
            j = seq[i]
            seq[i], seq[j] = seq[j], seq[i]
            s = -s
    return s


def _canonical(flav, pairs):
    """Maps each propagator (quark slot, antiquark slot) onto the canonical slot
    k of its flavour (flavours sorted up, dn, sr, ch; equal flavours in the
    order the pairs come).  Returns (slot of each pair, colour sign)."""
    order = sorted(range(3), key=lambda i: FLAVOURS.index(flav[pairs[i][0]]))
    free = [flav[pairs[i][0]] for i in order]           # canonical flavours
    slots = []
    taken = [False] * 3
    for q, a in pairs:
        k = next(k for k in range(3) if not taken[k] and free[k] == flav[q])
        taken[k] = True
        slots.append(k)
    # colour: quark slot q -> k, antiquark slot a -> 3 + k
    target = [0] * 6
    for (q, a), k in zip(pairs, slots):
        target[q] = k
        target[a] = 3 + k
    return slots, _parity(target), free


def baryon_terms(flavs, coeffs, spin2):
    """{operator: {(m0..m5): v}} and the canonical flavour triple of a baryon."""
    ops = OPS_HALF if spin2 == 1 else OPS_3HALF
    if spin2 == 3:
        norm = {3: 0.5, 2: 1.0, 1: 1.0 / 3.0}[len(coeffs)]
    else:
        norm = float(sum(c * c for c in coeffs))
    terms = []        # (coefficient, slots of the 6 fields' spins in the canonical tuple)
    canon = None
    for fj, cj in zip(flavs, coeffs):
        f1, f2, f3 = fj.split("-")
        for fb, cb in zip(flavs, coeffs):
            g1, g2, g3 = fb.split("-")
            flav = [f1, f2, f3, g3, g1, g2]        # q0 q1 q2 | qbar of Jbar: f3', f1', f2'
            for sign, pairs in _contractions(flav):
                slots, csign, free = _canonical(flav, pairs)
                canon = canon or free
                if free != canon:
                    raise ValueError("flavour content differs between terms")
                spin_pos = [0] * 6
                for (q, a), k in zip(pairs, slots):
                    spin_pos[q] = 2 * k
                    spin_pos[a] = 2 * k + 1
                terms.append((cj * cb * sign * csign, spin_pos))
    out = {}
    for name, triples in ops.items():
        acc = {}
        for P, G1, G2 in triples:
            # field spins: (q0, q1) by G1, (q2, qbar3) by the projector, (qbar4, qbar5) by G2
            for s in itertools.product(range(4), repeat=6):
                v = G1[s[0], s[1]] * P[s[2], s[3]] * G2[s[4], s[5]]
                if v == 0:
                    continue
                for c, pos in terms:
                    m = [0] * 6
                    for f in range(6):
                        m[pos[f]] = s[f]
                    m = tuple(m)
                    acc[m] = acc.get(m, 0) + c * v
        acc = {m: v / norm for m, v in acc.items() if abs(v) > 1e-8}
        # float32 coefficients, as stored by the reference
        out[name] = {m: complex(np.float32(v.real), np.float32(v.imag)) for m, v in acc.items()}
    return out, canon


def dataset_name(baryon, op):
    b = baryon.replace("+", "Pl").replace("-", "Mn").replace("*", "St").replace("'", "Pr")
    return b + "/" + op.replace("-", "_")


def products():
    """[(flavour letters 'udc', [(dataset name, {(m0..m5): v})])] in the
    reference's product order (first appearance in the baryon list)."""
    prods = {}
    for name, flavs, coeffs, spin2 in BARYONS:
        ops, canon = baryon_terms(flavs, coeffs, spin2)
        key = "".join(f[0] for f in canon)
        for op, t in ops.items():
            prods.setdefault(key, []).append((dataset_name(name, op), t))
    return list(prods.items())


def write_header(path):
    prods = products()
    lines = ["// Generated by plegma_amd/gen_udsc.py from the udsc baryon interpolating fields.",
             "// Per flavour product (propagators S1 S2 S3 in the order of its letters):",
             "// operators, each a list of terms {m0, m1, m2, m3, m4, m5, re, im}.",
             "// UdscTerm {m[6], re, im} is declared in plg_internal.h.",
             "#pragma once", "namespace plg {",
             "struct UdscOp { const char *name; int first, count; };",
             "struct UdscProd { const char *flavours; int first_op, nops; };"]
    terms, ops, prs = [], [], []
    for key, oplist in prods:
        prs.append(f'{{"{key}", {len(ops)}, {len(oplist)}}}')
        for name, t in oplist:
            ops.append(f'{{"{name}", {len(terms)}, {len(t)}}}')
            for m in sorted(t):
                v = t[m]
                terms.append("{{" + ",".join(str(x) for x in m) + "}," +
                             f"{float(np.float32(v.real))!r}f,{float(np.float32(v.imag))!r}f}}")
    lines.append(f"constexpr int UDSC_NPROD = {len(prs)}, UDSC_NOPS = {len(ops)}, UDSC_NTERMS = {len(terms)};")
    lines.append("static const UdscProd UDSC_PRODS[UDSC_NPROD] = {" + ", ".join(prs) + "};")
    lines.append("static const UdscOp UDSC_OPS[UDSC_NOPS] = {\n  " + ",\n  ".join(ops) + "};")
    body = ",".join(terms)
    lines.append("static const UdscTerm UDSC_TERMS[UDSC_NTERMS] = {" + body + "};")
    lines.append("}  // namespace plg")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    write_header(sys.argv[1])
