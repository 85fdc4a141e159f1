"""Writes udsc_golden.npz: the udsc baryon operators of the reference evaluated
on seeded site propagators.

The reference's tables (lib/kernels/PLEGMA_baryons_udsc_arrays.cuh:466-470 for
the product list, dataset names and counts; PLEGMA_baryons_udsc_arrays.cu for
the float2 coefficients and the six spin indices of every term) are read as
text, data only, and evaluated here with the colour structure of
lib/kernels/PLEGMA_baryons_udsc.cu:98-121:
    O = sum_t v_t sum_{eps, eps'} sgn sgn' S1[m0][m1][a][a'] S2[m2][m3][b][b'] S3[m4][m5][c][c']
on two sites of random propagators per flavour u, d, s, c.  The fixture holds
the propagators, the names and the values; the tables themselves are not kept.

Run from the repository root:  python tests/golden/make_udsc_golden.py /path/to/reference
"""
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

EPS = np.zeros((3, 3, 3))
for (a, b, c), s in {(0, 1, 2): 1, (1, 2, 0): 1, (2, 0, 1): 1,
                     (2, 1, 0): -1, (0, 2, 1): -1, (1, 0, 2): -1}.items():
    EPS[a, b, c] = s


def _strings(text, name):
    body = text[text.index(name + " = {") + len(name) + 4:]
    body = body[:body.index("};")]
    return body


def parse(ref):
    kd = os.path.join(ref, "lib", "kernels")
    head = open(os.path.join(kd, "PLEGMA_baryons_udsc_arrays.cuh")).read()
    data = open(os.path.join(kd, "PLEGMA_baryons_udsc_arrays.cu")).read()
    prods = re.findall(r'"(\w+)"', _strings(head, "BP_prop_prods"))
    names = [re.findall(r'"([^"]+)"', g) for g in
             re.findall(r"\{([^{}]*)\}", _strings(head, "BP_prop_prods_names"))]
    idx_names = [re.findall(r"(\w+)", g) for g in
                 re.findall(r"\{([^{}]*)\}", _strings(head, "BP_prop_prods_idxs"))]
    out = []
    for p, ns, ins in zip(prods, names, idx_names):
        ops = []
        for n, iname in zip(ns, ins):
            vname = iname[:-len("_idxs")] + "_vals"
            m = re.search(r"float2 " + vname + r"\[\d+\] = \{(.*?)\};", data)
            vals = [complex(np.float32(float(a)), np.float32(float(b)))
                    for a, b in re.findall(r"\{([^,{}]+), ([^,{}]+)\}", m.group(1))]
            m = re.search(r"short " + iname + r"\[\d+\*6\] = \{(.*?)\};", data)
            idx = np.array([int(x) for x in re.findall(r"-?\d+", m.group(1))]).reshape(-1, 6)
            assert len(idx) == len(vals)
            ops.append((n, idx, np.array(vals)))
        out.append((p, ops))
    return out


def evaluate(S, letters, idx, vals):
    """S: flavour -> [site, 4, 4, 3, 3]; returns [site]"""
    A = S[letters[0]][:, idx[:, 0], idx[:, 1]]
    B = S[letters[1]][:, idx[:, 2], idx[:, 3]]
    Cm = S[letters[2]][:, idx[:, 4], idx[:, 5]]
    t = np.einsum("abc,def,snad,snbe,sncf->sn", EPS, EPS, A, B, Cm, optimize=True)
    return t @ vals


def main(ref):
    rng = np.random.default_rng(20261202)
    S = {f: rng.standard_normal((2, 4, 4, 3, 3)) + 1j * rng.standard_normal((2, 4, 4, 3, 3)) for f in "udsc"}
    prods, names, values = [], [], []
    for p, ops in parse(ref):
        prods.append(p)
        for n, idx, vals in ops:
            names.append(p + ":" + n)
            values.append(evaluate(S, p, idx, vals))
    np.savez_compressed(os.path.join(HERE, "udsc_golden.npz"), props=np.stack([S[f] for f in "udsc"]),
                        products=np.array(prods), names=np.array(names), values=np.array(values))
    print(len(prods), "products,", len(names), "operators")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
