"""Writes tetra_golden.npz and tetra_bcud_golden.npz: the reference's
heavy-light and bcud tetraquark operators evaluated on seeded site
propagators, and a fingerprint of each table.

The reference's tables (lib/kernels/PLEGMA_heavy_light_tetraquarks_arrays.cuh:
124-129 for the product list, dataset names and counts;
PLEGMA_heavy_light_tetraquarks_arrays.cu for the float2 coefficients, the
eight spin indices and the eight colour slots of every term) are read as text,
data only (the bcud tables likewise from
lib/kernels/PLEGMA_bcud_tetraquarks_arrays.cuh:199-204 and
PLEGMA_bcud_tetraquarks_arrays.cu, the same format), and evaluated here with
the colour sums of
lib/kernels/PLEGMA_heavy_light_tetraquarks.cu:29-47:
    O = sum_t v_t sum_{a0..a3} P1[m0][m1][a_c0][a_c1] P2[m2][m3][a_c2][a_c3]
            conj(P3[(m5+2)%4][(m4+2)%4][a_c5][a_c4]) conj(P4[(m7+2)%4][(m6+2)%4][a_c7][a_c6])
on two sites of random propagators per flavour u, d, s, b (c for bcud; a
product "udbb" takes P1..P4 from its letters).  The fixture holds the propagators, the
names, the counts, a SHA-256 fingerprint of each table (int16 spins, int16
colour slots, float32 coefficient pairs, in the reference's term order) and the
values; the tables themselves are not kept.

Run from the repository root:  python tests/golden/make_tetra_golden.py /path/to/reference
"""
import hashlib
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _strings(text, name):
    body = text[text.index(name + " = {") + len(name) + 4:]
    return body[:body.index("};")]


def parse(ref, stem="PLEGMA_heavy_light_tetraquarks_arrays", var="TETRA_prop_prods"):
    kd = os.path.join(ref, "lib", "kernels")
    head = open(os.path.join(kd, stem + ".cuh")).read()
    data = open(os.path.join(kd, stem + ".cu")).read()
    prods = re.findall(r'"(\w+)"', _strings(head, var))
    names = [re.findall(r'"([^"]+)"', g) for g in
             re.findall(r"\{([^{}]*)\}", _strings(head, var + "_names"))]
    idx_names = [re.findall(r"(\w+)", g) for g in
                 re.findall(r"\{([^{}]*)\}", _strings(head, var + "_idxs"))]
    out = []
    for p, ns, ins in zip(prods, names, idx_names):
        ops = []
        for n, iname in zip(ns, ins):
            base = iname[:-len("_idxs")]
            m = re.search(r"float2 " + base + r"_vals\[\d+\] = \{(.*?)\};", data)
            vals = np.array([(np.float32(float(a)), np.float32(float(b)))
                             for a, b in re.findall(r"\{([^,{}]+), ([^,{}]+)\}", m.group(1))], dtype=np.float32)
            m = re.search(r"short " + iname + r"\[\d+\*8\] = \{(.*?)\};", data)
            idx = np.array([int(x) for x in re.findall(r"-?\d+", m.group(1))], dtype=np.int16).reshape(-1, 8)
            m = re.search(r"short " + base + r"_col_contr\[\d+\*8\] = \{(.*?)\};", data)
            col = np.array([int(x) for x in re.findall(r"-?\d+", m.group(1))], dtype=np.int16).reshape(-1, 8)
            assert len(idx) == len(vals) == len(col)
            ops.append((n, idx, col, vals))
        out.append((p, ops))
    return out


def fingerprint(idx, col, vals):
    """SHA-256 of a table in the reference's term order (a coefficient of -0
    counts as 0)"""
    h = hashlib.sha256()
    v = np.asarray(vals, dtype="<f4") + np.float32(0)
    for a in (np.asarray(idx, dtype="<i2"), np.asarray(col, dtype="<i2"), v):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def evaluate(S, letters, idx, col, vals):
    """S: flavour -> [site, 4, 4, 3, 3]; returns [site]"""
    A = np.array(np.meshgrid(*[np.arange(3)] * 4, indexing="ij")).reshape(4, -1).T   # [81][4]
    c = col.astype(int)
    sp = idx.astype(int)
    col_of = lambda j: A[:, c[:, j]].T                                               # [term][81]
    P = [S[f] for f in letters]
    t1 = P[0][:, sp[:, 0, None], sp[:, 1, None], col_of(0), col_of(1)]
    t2 = P[1][:, sp[:, 2, None], sp[:, 3, None], col_of(2), col_of(3)]
    t3 = np.conj(P[2][:, (sp[:, 5, None] + 2) % 4, (sp[:, 4, None] + 2) % 4, col_of(5), col_of(4)])
    t4 = np.conj(P[3][:, (sp[:, 7, None] + 2) % 4, (sp[:, 6, None] + 2) % 4, col_of(7), col_of(6)])
    v = vals[:, 0].astype(np.float64) + 1j * vals[:, 1].astype(np.float64)
    return (t1 * t2 * t3 * t4).sum(-1) @ v


def _write(path, tables, S, letters):
    prods, names, counts, prints, values = [], [], [], [], []
    for p, ops in tables:
        prods.append(p)
        for n, idx, col, vals in ops:
            names.append(p + ":" + n)
            counts.append(len(idx))
            prints.append(fingerprint(idx, col, vals))
            values.append(evaluate(S, p, idx, col, vals))
    np.savez_compressed(path, props=np.stack([S[f] for f in letters]), products=np.array(prods),
                        names=np.array(names), counts=np.array(counts), fingerprints=np.array(prints),
                        values=np.array(values))
    print(os.path.basename(path), len(prods), "products,", len(names), "operators")


def main(ref):
    rng = np.random.default_rng(20261203)
    S = {f: rng.standard_normal((2, 4, 4, 3, 3)) + 1j * rng.standard_normal((2, 4, 4, 3, 3)) for f in "udsb"}
    _write(os.path.join(HERE, "tetra_golden.npz"), parse(ref), S, "udsb")
    rng = np.random.default_rng(20261204)
    S = {f: rng.standard_normal((2, 4, 4, 3, 3)) + 1j * rng.standard_normal((2, 4, 4, 3, 3)) for f in "udcb"}
    _write(os.path.join(HERE, "tetra_bcud_golden.npz"),
           parse(ref, "PLEGMA_bcud_tetraquarks_arrays", "TETRA_bcud_prop_prods"), S, "udcb")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
