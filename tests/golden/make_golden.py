"""Writes the committed golden fixtures of this directory.

1. ref_tables.json -- the index/sign tables of the reference kernels, read as
   text (data only) from the reference sources:
     lib/kernels/PLEGMA_mesons.cuh:6-8        mesons_indices / mesons_values
     lib/kernels/PLEGMA_baryons_NtoN.cu:3-4   NtoN_indices / NtoN_values
     lib/kernels/PLEGMA_baryons_deltas.cu:3-4 deltas_indices / deltas_values
     lib/kernels/PLEGMA_gammas.cuh:6-41       gamma / gammaInd
   Only written when the reference tree is present (pass its path as argv[1]).
2. tiny_*.npz -- seeded propagators on 2^3x4 and the correlators computed from
   them by the independent closed forms in tests/closed_forms.py (Wick
   engine, spin traces, exact-phase DFT).  These pin the oracle.

Run from the repository root:  python tests/golden/make_golden.py [/path/to/reference]
"""
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import closed_forms as cf  # noqa: E402


def _array(text, name):
    text = re.sub(r"//[^\n]*", "", text)
    i = text.index(name)
    j = text.index("{", i)
    k = text.index("};", j)
    return [int(float(x)) for x in re.findall(r"-?\d+(?:\.\d+)?", text[j:k])]


def ref_tables(ref_root):
    kd = os.path.join(ref_root, "lib", "kernels")
    rd = lambda f: open(os.path.join(kd, f)).read()  # noqa: E731
    mes = rd("PLEGMA_mesons.cuh")
    nton = rd("PLEGMA_baryons_NtoN.cu")
    dl = rd("PLEGMA_baryons_deltas.cu")
    gm = rd("PLEGMA_gammas.cuh")
    return {
        "source": "PLEGMA lib/kernels (tables read as data)",
        "mesons_indices": _array(mes, "mesons_indices"),
        "mesons_values": _array(mes, "mesons_values"),
        "NtoN_indices": _array(nton, "NtoN_indices"),
        "NtoN_values": _array(nton, "NtoN_values"),
        "deltas_indices": _array(dl, "deltas_indices"),
        "deltas_values": _array(dl, "deltas_values"),
        "gamma": _array(gm, "float gamma[16]"),
        "gammaInd": _array(gm, "gammaInd[16]"),
    }


def philox_normal_props(L, seed, nprops):
    """Seeded complex N(0,1) propagators drawn with numpy (fixture inputs only)."""
    rng = np.random.default_rng(seed)
    V = L[0] * L[1] * L[2] * L[3]
    out = []
    for _ in range(nprops):
        z = rng.normal(size=(144, V)) + 1j * rng.normal(size=(144, V))
        out.append(z)
    return out


def sites(prop, L, t):
    V3 = L[0] * L[1] * L[2]
    return np.ascontiguousarray(prop[:, t * V3:(t + 1) * V3].T).reshape(V3, 4, 4, 3, 3)


def window(L, src, totalT):
    # lattice times in buffer-row order, derived from the covered interval:
    # the rows hold t0..t0+n-1 (mod T) with the part that wrapped past T first.
    T = L[3]
    n = min(totalT, T)
    times = [(src[3] + k) % T for k in range(n)]
    wrapped = [t for t in times if t < src[3]]
    direct = [t for t in times if t >= src[3]]
    return wrapped + direct


def correlators(L, src, totalT, q2max, S1, S2):
    moms = cf.momlist(q2max)
    rows = window(L, src, totalT)
    out = {k: [] for k in ("mesons10", "mesons16", "mesons_new", "mesons_all",
                           "baryons", "meson_meson")}
    for t in rows:
        A = sites(S1, L, t)
        B = sites(S2, L, t)
        m10 = np.stack([cf.meson(A, A, cf.GAMMAS[g]) for g in cf.MESON10] +
                       [cf.meson(B, B, cf.GAMMAS[g]) for g in cf.MESON10], -1)
        m16 = np.stack([cf.meson(A, A, G) for G in cf.GAMMAS] +
                       [cf.meson(B, B, G) for G in cf.GAMMAS], -1)
        mn = np.stack([cf.meson(A, B, cf.GAMMAS[g]) for g in cf.MESON10], -1)
        ma = [cf.meson(A, B, cf.GAMMAS[5]), cf.meson(A, B, cf.GAMMAS[0])]
        for p in range(2):
            for i in range(4):
                for j in range(4):
                    ma.append(cf.mesons_all_pair(A, B, cf.GAMMAS[1 + i + 5 * p],
                                                 cf.GAMMAS[1 + j + 5 * p]))
        ma = np.stack(ma, -1)
        # baryon groups nucl_nucl, deltapp_deltamm_11/22/33 ; datasets (S1,S2) roles
        bar = []
        for d, (u, v) in enumerate(((A, B), (B, A))):
            g = [cf.proton_wick(u, v).reshape(-1, 16)]
            for k in range(3):
                g.append(cf.delta32(u, k).reshape(-1, 16))
            bar.append(np.stack(g, 1))              # [V3, group, 16]
        bar = np.stack(bar, 1)                      # [V3, d, group, 16]
        mm_m, mm_mm = cf.meson_meson(A, B)
        mm = np.concatenate([mm_m, mm_mm], -1)
        for key, f in (("mesons10", m10), ("mesons16", m16), ("mesons_new", mn),
                       ("mesons_all", ma), ("baryons", bar), ("meson_meson", mm)):
            out[key].append(cf.dft(f, L, src[:3], moms))
    res = {k: np.stack(v, 0) for k, v in out.items()}   # [row, mom, ...]
    return res, np.array(rows), moms


def main():
    if len(sys.argv) > 1 and os.path.isdir(sys.argv[1]):
        with open(os.path.join(HERE, "ref_tables.json"), "w") as f:
            json.dump(ref_tables(sys.argv[1]), f)
    L = [2, 2, 2, 4]
    S1, S2 = philox_normal_props(L, 20240601, 2)
    for name, src, totalT in (("tiny_full", [1, 0, 1, 2], 4), ("tiny_wrap", [0, 1, 1, 3], 3)):
        res, rows, moms = correlators(L, src, totalT, 3, S1, S2)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), L=np.array(L), src=np.array(src),
                            totalT=totalT, q2max=3, S1=S1, S2=S2, rows=rows, moms=moms, **res)
    print("golden fixtures written")


if __name__ == "__main__":
    main()
