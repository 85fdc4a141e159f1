"""Four-propagator table contractions (the tetraquark table format of
lib/kernels/PLEGMA_heavy_light_tetraquarks.cu:26-49 and
lib/kernels/PLEGMA_bcud_tetraquarks.cu:29-47): the oracle's literal 81-term
colour sums against the closed forms the colour slots close into, for the
colour patterns the reference's tables use; the generated heavy-light and
bcud tables against the reference's."""
import numpy as np
import pytest

from oracle import oracle as O

# (c0..c7) of the reference's tables: two close as <B1,B3><B2,B4>, two as
# Tr[(B3^dag B1)(B4^dag B2)]
DIRECT = [(0, 2, 1, 3, 2, 0, 3, 1), (0, 3, 1, 2, 3, 0, 2, 1)]
CROSSED = [(0, 2, 1, 3, 3, 0, 2, 1), (0, 3, 1, 2, 2, 0, 3, 1)]
# the bcud tables' eight patterns: four pair P1 with P3 and P2 with P4 as
# above, four pair P1 with P4 and P2 with P3 (<B1,B4><B2,B3> and
# Tr[(B4^dag B1)(B3^dag B2)])
BCUD_DIRECT = [(1, 3, 0, 2, 3, 1, 2, 0), (1, 2, 0, 3, 2, 1, 3, 0)]
BCUD_CROSSED = [(1, 2, 0, 3, 3, 1, 2, 0), (1, 3, 0, 2, 2, 1, 3, 0)]
SWAP_DIRECT = [(0, 2, 1, 3, 3, 1, 2, 0), (0, 3, 1, 2, 2, 1, 3, 0)]
SWAP_CROSSED = [(0, 3, 1, 2, 3, 1, 2, 0), (0, 2, 1, 3, 2, 1, 3, 0)]


def _closed(S, m, form):
    B1 = S[0][m[0], m[1]]
    B2 = S[1][m[2], m[3]]
    B3 = S[2][(m[5] + 2) % 4, (m[4] + 2) % 4]
    B4 = S[3][(m[7] + 2) % 4, (m[6] + 2) % 4]
    if form == "direct":
        return np.sum(B1 * B3.conj()) * np.sum(B2 * B4.conj())
    if form == "crossed":
        return np.trace((B3.conj().T @ B1) @ (B4.conj().T @ B2))
    if form == "swap-direct":
        return np.sum(B1 * B4.conj()) * np.sum(B2 * B3.conj())
    return np.trace((B4.conj().T @ B1) @ (B3.conj().T @ B2))


@pytest.mark.parametrize("cols,form", [(c, "direct") for c in DIRECT + BCUD_DIRECT] +
                         [(c, "crossed") for c in CROSSED + BCUD_CROSSED] +
                         [(c, "swap-direct") for c in SWAP_DIRECT] + [(c, "swap-crossed") for c in SWAP_CROSSED])
def test_colour_patterns_close(cols, form):
    rng = np.random.default_rng(3)
    S = [rng.standard_normal((4, 4, 3, 3)) + 1j * rng.standard_normal((4, 4, 3, 3)) for _ in range(4)]
    raw = [np.ascontiguousarray(np.stack([s.real, s.imag], -1).reshape(-1)) for s in S]
    for _ in range(6):
        m = rng.integers(0, 4, size=8)
        v = complex(rng.standard_normal(), rng.standard_normal())
        idx = np.array(m, dtype=np.int32)
        col = np.array(cols, dtype=np.int32)
        val = np.array([v.real, v.imag])
        out = np.zeros(2)
        ip = O.C.POINTER(O.C.c_int)
        O.lib().orc_site_tetra(*[O._ptr(r) for r in raw], idx.ctypes.data_as(ip), col.ctypes.data_as(ip),
                               O._ptr(val), np.array([0], dtype=np.int32).ctypes.data_as(ip),
                               np.array([1], dtype=np.int32).ctypes.data_as(ip), 1, O._ptr(out))
        got = complex(out[0], out[1])
        ref = v * _closed(S, m, form)
        assert abs(got - ref) <= 1e-12 * abs(ref)


def test_unit_propagators():
    """S = 1 (spin x colour): a direct term is 3 x 3 = 9 on the spin diagonal,
    a crossed one Tr[1] = 3; off-diagonal spins give 0"""
    S = np.zeros((4, 4, 3, 3), dtype=complex)
    for mu in range(4):
        S[mu, mu] = np.eye(3)
    raw = np.ascontiguousarray(np.stack([S.real, S.imag], -1).reshape(-1))
    ip = O.C.POINTER(O.C.c_int)
    for cols, want in ((DIRECT[0], 9), (CROSSED[1], 3)):
        for m, on in (((0, 0, 1, 1, 2, 2, 3, 3), True), ((0, 1, 1, 1, 2, 2, 3, 3), False)):
            out = np.zeros(2)
            O.lib().orc_site_tetra(*[O._ptr(raw)] * 4, np.array(m, dtype=np.int32).ctypes.data_as(ip),
                                   np.array(cols, dtype=np.int32).ctypes.data_as(ip), O._ptr(np.array([1.0, 0.0])),
                                   np.array([0], dtype=np.int32).ctypes.data_as(ip),
                                   np.array([1], dtype=np.int32).ctypes.data_as(ip), 1, O._ptr(out))
            assert complex(out[0], out[1]) == (want if on else 0)


def _site_ops(S, idx, col, vals, counts):
    """orc_site_tetra on one site: all operators of a table"""
    raw = [np.ascontiguousarray(np.stack([s.real, s.imag], -1).reshape(-1)) for s in S]
    ip = O.C.POINTER(O.C.c_int)
    idx = np.ascontiguousarray(np.asarray(idx, dtype=np.int32).reshape(-1))
    col = np.ascontiguousarray(np.asarray(col, dtype=np.int32).reshape(-1))
    val = np.ascontiguousarray(np.stack([vals.real, vals.imag], -1).astype(np.float64).reshape(-1))
    cnt = np.asarray(counts, dtype=np.int32)
    first = np.concatenate([[0], np.cumsum(cnt)[:-1]]).astype(np.int32)
    out = np.zeros(2 * len(cnt))
    O.lib().orc_site_tetra(*[O._ptr(r) for r in raw], idx.ctypes.data_as(ip), col.ctypes.data_as(ip), O._ptr(val),
                           first.ctypes.data_as(ip), cnt.ctypes.data_as(ip), len(cnt), O._ptr(out))
    return out[0::2] + 1j * out[1::2]


def test_generated_tables_equal_the_reference():
    """plegma_amd/gen_tetra.py derives the heavy-light tetraquark tables from
    the interpolating fields; the golden fixture holds the SHA-256 of every
    reference table (spins, colour slots, float32 coefficients, term order),
    so equal fingerprints mean bit-identical tables"""
    import os
    from plegma_amd import gen_tetra
    from golden.make_tetra_golden import fingerprint
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "tetra_golden.npz"))
    tabs = gen_tetra.tables()
    assert [p for p, _ in tabs] == list(g["products"])
    names = [p + ":" + n for p, ops in tabs for n, *_ in ops]
    assert names == list(g["names"])
    counts = [len(v) for _, ops in tabs for *_, v in ops]
    assert counts == list(g["counts"])
    prints = [fingerprint(i, c, np.stack([v.real, v.imag], -1)) for _, ops in tabs for _, i, c, v in ops]
    assert prints == list(g["fingerprints"])


def test_oracle_on_generated_tables_matches_golden():
    """the oracle's literal 81-term sums over the generated tables reproduce
    the reference tables' values (evaluated independently when the fixture
    was made) on both golden sites"""
    import os
    from plegma_amd import gen_tetra
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "tetra_golden.npz"))
    props = dict(zip("udsb", g["props"]))
    k = 0
    for p, ops in gen_tetra.tables():
        idx = np.concatenate([o[1] for o in ops])
        col = np.concatenate([o[2] for o in ops])
        val = np.concatenate([o[3] for o in ops]).astype(np.complex128)
        cnt = [len(o[3]) for o in ops]
        for site in range(2):
            got = _site_ops([props[f][site] for f in p], idx, col, val, cnt)
            ref = g["values"][k:k + len(ops), site]
            np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
        k += len(ops)


def test_bcud_tables_use_eight_patterns():
    """every term of the generated bcud tables has one of the eight patterns
    whose closed forms are checked above"""
    from plegma_amd import gen_tetra
    known = set(DIRECT + CROSSED + BCUD_DIRECT + BCUD_CROSSED + SWAP_DIRECT + SWAP_CROSSED)
    (_, ops), = gen_tetra.bcud_tables()
    seen = {tuple(int(x) for x in c) for o in ops for c in o[2]}
    assert seen <= known
    assert seen & set(SWAP_DIRECT + SWAP_CROSSED) and seen & set(BCUD_DIRECT + BCUD_CROSSED)


def test_generated_bcud_tables_equal_the_reference():
    """plegma_amd/gen_tetra.py derives the bcud tables (J = 1 and J = 0, the
    flavour combination (bbar u cbar d) - (bbar d cbar u)) from the
    interpolating fields; the golden fixture holds the SHA-256 of each of the
    reference's 57 tables"""
    import os
    from plegma_amd import gen_tetra
    from golden.make_tetra_golden import fingerprint
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "tetra_bcud_golden.npz"))
    tabs = gen_tetra.bcud_tables()
    assert [p for p, _ in tabs] == list(g["products"]) == ["udcb"]
    assert [p + ":" + n for p, ops in tabs for n, *_ in ops] == list(g["names"])
    assert [len(v) for _, ops in tabs for *_, v in ops] == list(g["counts"])
    prints = [fingerprint(i, c, np.stack([v.real, v.imag], -1)) for _, ops in tabs for _, i, c, v in ops]
    assert prints == list(g["fingerprints"])


def test_oracle_on_bcud_tables_matches_golden():
    """the oracle's 81-term sums over the generated bcud tables reproduce the
    reference tables' values on both golden sites"""
    import os
    from plegma_amd import gen_tetra
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "tetra_bcud_golden.npz"))
    props = dict(zip("udcb", g["props"]))
    (p, ops), = gen_tetra.bcud_tables()
    idx = np.concatenate([o[1] for o in ops])
    col = np.concatenate([o[2] for o in ops])
    val = np.concatenate([o[3] for o in ops]).astype(np.complex128)
    cnt = [len(o[3]) for o in ops]
    for site in range(2):
        got = _site_ops([props[f][site] for f in p], idx, col, val, cnt)
        ref = g["values"][:, site]
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
