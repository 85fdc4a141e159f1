"""Independent closed forms of the 2pt contractions (numpy, FP64).

These are written from the physics formulas (spin traces, the full epsilon
tensor, Wick terms as 12x12 products), not from the reference's per-term
tables, so they pin the oracle's table-driven restatement of
lib/kernels/PLEGMA_mesons.cuh, PLEGMA_baryons_NtoN.cu and
PLEGMA_baryons_deltas.cu against a second derivation.
"""
import numpy as np

# gamma basis of lib/kernels/PLEGMA_gammas.cuh:6-41 (Appendix: chiral-like basis)
g1 = np.array([[0, 0, 0, 1j], [0, 0, 1j, 0], [0, -1j, 0, 0], [-1j, 0, 0, 0]])
g2 = np.array([[0, 0, 0, 1], [0, 0, -1, 0], [0, -1, 0, 0], [1, 0, 0, 0]], dtype=complex)
g3 = np.array([[0, 0, 1j, 0], [0, 0, 0, -1j], [-1j, 0, 0, 0], [0, 1j, 0, 0]])
g4 = np.diag([1, 1, -1, -1]).astype(complex)
g5 = g1 @ g2 @ g3 @ g4
ONE = np.eye(4, dtype=complex)


def _sig(a, b):
    return -0.5j * (a @ b - b @ a)


# GAMMAS order, include/global/PLEGMA_enums.h:22
GAMMAS = [ONE, g1, g2, g3, g4, g5, g5 @ g1, g5 @ g2, g5 @ g3, g5 @ g4,
          _sig(g1, g2), _sig(g1, g3), _sig(g2, g3), _sig(g4, g1), _sig(g4, g2), _sig(g4, g3)]
MESON10 = [5, 0, 6, 7, 8, 9, 1, 2, 3, 4]
Cc = 1j * g2 @ g4

EPS = np.zeros((3, 3, 3))
for (a, b, c), s in {(0, 1, 2): 1, (1, 2, 0): 1, (2, 0, 1): 1,
                     (2, 1, 0): -1, (0, 2, 1): -1, (1, 0, 2): -1}.items():
    EPS[a, b, c] = s


def bar(A):
    return g4 @ A.conj().T @ g4


def meson(S1, S2, G):
    """sum_ab Tr_s[G S1^ab g5 G (S2^ab)^dag g5]; S[..., mu, nu, a, b] arrays."""
    X = g5 @ G
    # Tr[G S1 X S2^dag g5] over spin, colour traced pairwise
    return np.einsum('ia,...abxy,bc,...dcxy,di->...', G, S1, X, S2.conj(), g5)


def mesons_all_pair(S1, S2, Gm, Gn):
    """sum_ab Tr_s[Gm S1^ab Gn (S2^ab)^dag]"""
    return np.einsum('da,...abxy,bz,...dzxy->...', Gm, S1, Gn, S2.conj())


def proton(S1, S2, A=None):
    """spin-open proton, C_{mu mu'}: eps eps' A_{nu lam} Abar_{lam' nu'}
    S2^{bb'}_{lam lam'} [S1^{aa'}_{nu nu'} S1^{cc'}_{mu mu'} - S1^{ac'}_{nu mu'} S1^{ca'}_{mu nu'}]"""
    if A is None:
        A = Cc @ g5
    Ab = bar(A)
    # S[..., spin, spin, col, col]
    t1 = np.einsum('abc,xyz,nl,pq,...lpby,...nqax,...mMcz->...mM',
                   EPS, EPS, A, Ab, S2, S1, S1, optimize=True)
    t2 = np.einsum('abc,xyz,nl,pq,...lpby,...nMaz,...mqcx->...mM',
                   EPS, EPS, A, Ab, S2, S1, S1, optimize=True)
    return t1 - t2


def _perm_sign(p):
    p = list(p)
    s = 1
    for i in range(len(p)):
        while p[i] != i:
            j = p[i]
            p[i], p[j] = p[j], p[i]
            s = -s
    return s


def baryon_wick(props, sink, source, A, Ab=None):
    """Spin-open baryon 2pt from Wick's theorem.

    Interpolator chi_mu = sum_terms w * eps_abc (q1_{nu a} A_{nu lam} q2_{lam b}) q3_{mu c}
    with flavours (q1,q2,q3) per term; the conjugate uses Abar = g4 A^dag g4
    (or the given source matrix Ab).
    Every pairing of sink quark i with a source antiquark sigma(i) of the same
    flavour contributes sign(sigma) * prod_i S_f[s_i][s'_sigma(i)][c_i][c'_sigma(i)].
    props: dict flavour -> S[..., 4,4,3,3]."""
    import itertools
    if Ab is None:
        Ab = bar(A)
    sp, spp = 'nlm', 'qpM'      # sink spins (nu,lam,mu), source spins (nu',lam',mu')
    co, cop = 'abc', 'xyz'
    total = 0
    for ws, fs in sink:
        for wt, ft in source:
            for sig in itertools.permutations(range(3)):
                if any(fs[i] != ft[sig[i]] for i in range(3)):
                    continue
                specs = ['...' + sp[i] + spp[sig[i]] + co[i] + cop[sig[i]] for i in range(3)]
                term = np.einsum('abc,xyz,nl,pq,' + ','.join(specs) + '->...mM', EPS, EPS, A, Ab,
                                 *[props[fs[i]] for i in range(3)], optimize=True)
                total = total + ws * wt * _perm_sign(sig) * term
    return total


def proton_wick(S1, S2):
    """uud proton, Cg5 diquark: dataset 1 of contractBaryons (u=S1, d=S2)."""
    return baryon_wick({'u': S1, 'd': S2}, [(1, 'udu')], [(1, 'udu')], Cc @ g5)


def delta32(S, k):
    """Delta++ (uuu), Cg_k diquark."""
    return baryon_wick({'u': S}, [(1, 'uuu')], [(1, 'uuu')], Cc @ [g1, g2, g3][k])


def delta12(S1, S2, k):
    """Delta+ (iso 1/2): (2 (u Cg_k d) u + (u Cg_k u) d) / sqrt(3), u=S1, d=S2."""
    t = [(2, 'udu'), (1, 'uud')]
    return baryon_wick({'u': S1, 'd': S2}, t, t, Cc @ [g1, g2, g3][k]) / 3.0


def to12(S):
    """[..., mu, nu, a, b] -> [..., 12, 12] with index spin*3+colour"""
    sh = S.shape[:-4]
    return S.transpose(*range(len(sh)), -4, -2, -3, -1).reshape(*sh, 12, 12)


def meson_meson(SL, SH):
    """meson(G)=Tr A_G (16), mesmes(G)=(Tr A)^2 - Tr[A A] for G in g5,g1,g2,g3,
    A_G = G SL g5 G SH^dag g5 on 12x12 spin-colour matrices."""
    L12, H12 = to12(SL), to12(SH)
    I3 = np.eye(3)
    out_m, out_mm = [], []
    for op, G in enumerate(GAMMAS):
        Gs = np.kron(G, I3)
        X = np.kron(g5 @ G, I3)
        G5 = np.kron(g5, I3)
        A = Gs @ L12 @ X @ np.conj(np.swapaxes(H12, -1, -2)) @ G5
        tr = np.trace(A, axis1=-2, axis2=-1)
        out_m.append(tr)
        if op in (5, 1, 2, 3):
            out_mm.append((op, tr * tr - np.trace(A @ A, axis1=-2, axis2=-1)))
    mm = [v for _, v in sorted(out_mm, key=lambda p: [5, 1, 2, 3].index(p[0]))]
    return np.stack(out_m, -1), np.stack(mm, -1)


def momlist(q2max):
    out = []
    for iQ in range(q2max + 1):
        for nx in range(iQ, -iQ - 1, -1):
            for ny in range(iQ, -iQ - 1, -1):
                for nz in range(iQ, -iQ - 1, -1):
                    if nx * nx + ny * ny + nz * nz == iQ:
                        out.append((nx, ny, nz))
    return np.array(out, dtype=np.int64).reshape(-1, 3)


def dft(f, L, src3, moms):
    """f: [V3, ...] in site order x + Lx(y + Ly z); exact phases via n.(x-x0) mod L."""
    V3 = L[0] * L[1] * L[2]
    x = np.arange(V3)
    c = np.stack([x % L[0], (x // L[0]) % L[1], x // (L[0] * L[1])], -1) - np.array(src3)
    # phase = 2 pi sum_k n_k d_k / L_k, computed with a common denominator
    M = int(np.lcm.reduce(L[:3]))
    k = (moms[:, None, :] * c[None, :, :] * (M // np.array(L[:3]))[None, None, :]).sum(-1) % M
    P = np.exp(-2j * np.pi * k / M)
    return np.tensordot(P, f.reshape(V3, -1), axes=(1, 0)).reshape(len(moms), *f.shape[1:])
