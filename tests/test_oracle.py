"""CPU tests pinning the oracle (oracle/) before it is trusted as the checker.

Pins: the reference's own index/sign tables (tests/golden/ref_tables.json,
read as data from the reference kernels), golden vectors from independent
closed forms (tests/golden/tiny_*.npz), Random123 known answers for Philox,
closed-form special cases and DFT invariants.
"""
import json
import os

import numpy as np
import pytest

import closed_forms as cf
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def ref():
    with open(os.path.join(GOLD, "ref_tables.json")) as f:
        return json.load(f)


def _terms(idx, val):
    return sorted((tuple(int(v) for v in i), complex(x)) for i, x in zip(idx, val))


def test_meson_tables_match_reference(ref):
    # PLEGMA_mesons.cuh:6-8: rows are (beta, gamma, delta, alpha), real values
    mi = np.array(ref["mesons_indices"]).reshape(10, 16, 4)
    mv = np.array(ref["mesons_values"]).reshape(10, 16)
    for ip, op in enumerate(cf.MESON10):
        idx, val = O.meson_table(op)
        assert _terms(idx, val) == _terms(mi[ip], mv[ip]), f"channel {ip}"


def test_diquark_tables_match_reference(ref):
    ni = np.array(ref["NtoN_indices"]).reshape(16, 4)
    nv = np.array(ref["NtoN_values"])
    idx, val = O.diquark_table(0)
    assert _terms(idx, val) == _terms(ni, nv)
    di = np.array(ref["deltas_indices"]).reshape(3, 16, 4)
    dv = np.array(ref["deltas_values"]).reshape(3, 16)
    for k in range(3):
        idx, val = O.diquark_table(k + 1)
        assert _terms(idx, val) == _terms(di[k], dv[k]), f"Cg{k+1}"


def test_gamma_basis_matches_reference(ref):
    gv = np.array(ref["gamma"]).reshape(16, 4, 2)
    gi = np.array(ref["gammaInd"]).reshape(16, 4, 2)
    for op in range(16):
        M = np.zeros((4, 4), complex)
        for nz in range(4):
            M[gi[op, nz, 0], gi[op, nz, 1]] = gv[op, nz, 0] + 1j * gv[op, nz, 1]
        np.testing.assert_allclose(O.gamma(op), M, atol=1e-15)
        np.testing.assert_allclose(cf.GAMMAS[op], M, atol=1e-15)


def test_clifford_and_hermiticity():
    g = [O.gamma(i) for i in (1, 2, 3, 4)]
    for i in range(4):
        for j in range(4):
            ac = g[i] @ g[j] + g[j] @ g[i]
            np.testing.assert_allclose(ac, 2 * np.eye(4) * (i == j), atol=1e-15)
        np.testing.assert_allclose(g[i], g[i].conj().T, atol=1e-15)
    np.testing.assert_allclose(O.gamma(5), g[0] @ g[1] @ g[2] @ g[3], atol=1e-15)


def test_philox_known_answers():
    # Random123 known-answer vectors for philox4x32_10
    assert O.philox([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert O.philox([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e,
                                                           0xa20bc7c6, 0x6d5451fd]
    assert O.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
                    [0xa4093822, 0x299f31d0]) == [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_normal_generator_moments():
    z = np.array([O.normal_pair(7, 0, s, c) for s in range(200) for c in range(144)])
    assert abs(z.mean()) < 0.02 and abs(z.std() - 1) < 0.02


@pytest.mark.parametrize("q,n", [(0, 1), (1, 7), (2, 19), (3, 27), (4, 33), (9, 123), (16, 257)])
def test_momentum_counts(q, n):
    assert len(O.momlist(q)) == n


def test_momentum_order():
    m = O.momlist(16)
    q = (m ** 2).sum(1)
    assert np.all(np.diff(q) >= 0)
    for s in np.unique(q):
        sh = [tuple(x) for x in m[q == s]]
        assert sh == sorted(sh, reverse=True)
    assert [tuple(x) for x in m[1:7]] == [(1, 0, 0), (0, 1, 0), (0, 0, 1), (0, 0, -1),
                                          (0, -1, 0), (-1, 0, 0)]
    np.testing.assert_array_equal(m, cf.momlist(16))


@pytest.mark.parametrize("T,t0,totalT", [(64, 10, 64), (64, 60, 16), (16, 0, 16), (16, 5, 8),
                                         (16, 15, 3), (8, 3, 8), (64, 0, 20), (64, 63, 32)])
def test_window_rows(T, t0, totalT):
    L = [4, 4, 4, T]
    src = [1, 2, 3, t0]
    s, e, lt = O.window(L, src, totalT)
    n = min(totalT, T)
    assert lt == n
    times = [O.row_time(L, src, totalT, j) for j in range(lt)]
    assert sorted(times) == sorted((t0 + k) % T for k in range(n))
    # file index (startT + j) mod totalT is the source-relative time
    for j, t in enumerate(times):
        assert (s + j) % totalT == (t - t0) % T


@pytest.mark.parametrize("name", ["tiny_full", "tiny_wrap"])
def test_oracle_matches_golden(name):
    g = np.load(os.path.join(GOLD, name + ".npz"))
    L = [int(x) for x in g["L"]]
    src = [int(x) for x in g["src"]]
    totalT, q2 = int(g["totalT"]), int(g["q2max"])
    soa = lambda z: np.ascontiguousarray(np.stack([z.real, z.imag], -1).reshape(-1))  # noqa
    P1, P2 = soa(g["S1"]), soa(g["S2"])
    np.testing.assert_array_equal(O.momlist(q2), g["moms"])
    rows = [O.row_time(L, src, totalT, j) for j in range(len(g["rows"]))]
    assert rows == list(g["rows"])

    def chk(kind, ref, mask=O.G_NUCL):
        H = O.correlator(kind, L, src, totalT, q2, P1, P2, mask=mask)
        err = np.abs(H - ref).max() / np.abs(ref).max()
        assert err < 1e-12, (kind, err)

    nr = len(rows)
    m10 = g["mesons10"]                                   # [row, mom, 20]
    chk(O.MESONS, m10.reshape(nr, -1, 2, 10).transpose(2, 0, 1, 3)[None])
    m16 = g["mesons16"]
    chk(O.MESONS16, m16.reshape(nr, -1, 2, 16).transpose(2, 0, 1, 3)[None])
    chk(O.MESONS_NEW, g["mesons_new"][None, None])
    chk(O.MESONS_ALL, g["mesons_all"][None, None])
    b = g["baryons"]                                      # [row, mom, d, group, 16]
    chk(O.BARYONS, b.transpose(3, 2, 0, 1, 4), mask=O.G_NUCL | sum(O.G_DELTA32))
    chk(O.MESON_MESON, g["meson_meson"][None, None])


def test_unit_propagator_closed_forms():
    S = np.einsum("mn,ab->mnab", np.eye(4), np.eye(3)).astype(complex)
    p = O.site_body("nton", S, S)
    np.testing.assert_allclose(p[:16].reshape(4, 4), 30 * np.eye(4), atol=1e-13)
    m = O.site_body("mesons16", S, S)
    for op in range(16):
        G = cf.GAMMAS[op]
        np.testing.assert_allclose(m[op], 3 * np.trace(G @ cf.g5 @ G @ cf.g5), atol=1e-13)
    for k in range(3):
        # uuu with S = 1: the Wick engine's six delta-function permutations
        d = O.site_body("delta32", S, S, k)[:16].reshape(4, 4)
        np.testing.assert_allclose(d, cf.delta32(S[None], k)[0], atol=1e-12)


def test_delta12_matches_wick_closed_form():
    """orc_site_delta12 (restating PLEGMA_baryons_deltas.cu:3-60: the 8-term
    iso-1/2 sum with weights -4, 2, 2, -2, -2, -1, 1, 4 over 3) against the
    Wick engine of (2 (u Cg_k d) u + (u Cg_k u) d) / sqrt(3), with u and d
    distinct random propagators, so that a dropped term, a wrong weight or a
    swapped flavour changes the result; dataset 2 swaps the flavours."""
    rng = np.random.default_rng(12)
    S1 = rng.normal(size=(4, 4, 3, 3)) + 1j * rng.normal(size=(4, 4, 3, 3))
    S2 = rng.normal(size=(4, 4, 3, 3)) + 1j * rng.normal(size=(4, 4, 3, 3))
    for k in range(3):
        d = O.site_body("delta12", S1, S2, k)
        ref1 = cf.delta12(S1[None], S2[None], k)[0]
        ref2 = cf.delta12(S2[None], S1[None], k)[0]
        scale = max(np.abs(ref1).max(), np.abs(ref2).max())
        np.testing.assert_allclose(d[:16].reshape(4, 4), ref1, atol=1e-12 * scale)
        np.testing.assert_allclose(d[16:].reshape(4, 4), ref2, atol=1e-12 * scale)
    # u = d: the iso-1/2 and iso-3/2 states are made of the same quarks, and
    # (2 udu + uud)/sqrt(3) with u = d is sqrt(3) (uuu), so Delta+ = 3 Delta++ / 3
    for k in range(3):
        d12 = O.site_body("delta12", S1, S1, k)[:16]
        d32 = O.site_body("delta32", S1, S1, k)[:16]
        np.testing.assert_allclose(d12, d32, atol=1e-12 * np.abs(d32).max())


def test_pion_positive():
    rng = np.random.default_rng(3)
    S = rng.normal(size=(4, 4, 3, 3)) + 1j * rng.normal(size=(4, 4, 3, 3))
    m = O.site_body("mesons10", S, S)
    np.testing.assert_allclose(m[0], np.sum(np.abs(S) ** 2), rtol=1e-14)


def test_dft_invariants():
    L = [4, 6, 2, 4]
    V3 = 48
    rng = np.random.default_rng(1)
    f = rng.normal(size=(V3, 3)) + 1j * rng.normal(size=(V3, 3))
    moms = O.momlist(5)
    src = [1, 4, 1]
    F = O.dft(L, src, 5, f)
    np.testing.assert_allclose(F[0], f.sum(0), rtol=1e-13)           # n = 0
    np.testing.assert_allclose(F, cf.dft(f, L, src, moms), atol=1e-12)
    # delta field at site x gives exp(-i 2pi n.(x-x0)/L)
    d = np.zeros((V3, 1), complex)
    d[17, 0] = 1
    x = np.array([17 % 4, (17 // 4) % 6, 17 // 24]) - src
    ph = np.exp(-2j * np.pi * (moms * x / np.array(L[:3])).sum(1))
    np.testing.assert_allclose(O.dft(L, src, 5, d)[:, 0], ph, atol=1e-14)
    # shift theorem: moving the source multiplies by exp(+i 2pi n.dx/L)
    F2 = O.dft(L, [0, 0, 0], 5, f)
    ph2 = np.exp(2j * np.pi * (moms * np.array(src) / np.array(L[:3])).sum(1))
    np.testing.assert_allclose(F, F2 * ph2[:, None], atol=1e-12)
