"""Propagator preparation in the CPU oracle (rotation to the physical basis,
boundary sign), pinned by closed forms:
- the rotation is S -> U S U with U = (1 + i s g5)/sqrt(2) in spin space
  (g5 of lib/kernels/PLEGMA_gammas.cuh, built independently in closed_forms);
- rotating with s and then -s is the identity (U(-s) U(s) = 1);
- the boundary fix negates exactly the planes with global t < t0, so meson
  correlators (bilinear) are unchanged and baryon correlators (trilinear) flip
  sign on those time slices.
"""
import numpy as np
import pytest

import closed_forms as cf
from oracle import oracle as O


@pytest.fixture(scope="module", autouse=True)
def _lib():
    O.build()


def _rand_site(seed):
    r = np.random.default_rng(seed)
    return r.normal(size=144) + 1j * r.normal(size=144)


@pytest.mark.parametrize("sign", [1, -1])
def test_rotation_is_u_s_u(sign):
    S = _rand_site(3 + sign)
    U = (np.eye(4) + 1j * sign * cf.g5) / np.sqrt(2)
    M = S.reshape(4, 4, 3, 3)                      # [mu][nu][a][b]
    ref = np.einsum("mk,knab,nl->mlab", U, M, U).reshape(144)
    got = O.rotate_site(S, sign)
    assert np.abs(got - ref).max() < 1e-14 * np.abs(ref).max()


def test_rotation_inverse():
    S = _rand_site(11)
    back = O.rotate_site(O.rotate_site(S, 1), -1)
    assert np.abs(back - S).max() < 1e-14


def test_prop_rotate_matches_site():
    L = [2, 2, 2, 4]
    V3 = 8
    r = np.random.default_rng(5)
    P = np.ascontiguousarray(r.normal(size=(144, L[3] * V3, 2)))
    Q = P.copy()
    O.prop_rotate(Q, L, -1)
    for e in (0, 7, 13, 31):
        S = P[:, e, 0] + 1j * P[:, e, 1]
        ref = O.rotate_site(S, -1)
        assert np.array_equal(Q[:, e, 0] + 1j * Q[:, e, 1], ref)


@pytest.mark.parametrize("t_first,t0", [(0, 3), (5, 2), (6, 7)])
def test_boundaries_negate_planes_before_t0(t_first, t0):
    L = [2, 2, 2, 8]
    V3 = 8
    nt = 4
    r = np.random.default_rng(t0)
    P = np.ascontiguousarray(r.normal(size=(144, nt * V3, 2)))
    Q = P.copy()
    O.prop_boundaries(Q, L, t0, t_first)
    for pl in range(nt):
        t = (t_first + pl) % L[3]
        blk = slice(pl * V3, (pl + 1) * V3)
        assert np.array_equal(Q[:, blk], -P[:, blk] if t < t0 else P[:, blk])


def test_boundaries_flip_baryons_not_mesons():
    L = [2, 2, 2, 4]
    P = O.synth_prop(L, 99, 1, 0, 0.5)
    Pb = np.ascontiguousarray(P.reshape(144, -1, 2).copy())
    t0 = 2
    O.prop_boundaries(Pb, L, t0)
    Pb = Pb.reshape(-1)
    src, T = [0, 0, 0, 0], L[3]
    for kind, mask, flips in ((O.MESONS16, 0, False), (O.BARYONS, O.G_NUCL, True)):
        a = O.correlator(kind, L, src, T, 1, P, P, mask=mask)    # [group][dataset][t][mom][ch]
        b = O.correlator(kind, L, src, T, 1, Pb, Pb, mask=mask)
        for t in range(T):
            expect = -a[:, :, t] if (flips and t < t0) else a[:, :, t]
            assert np.abs(b[:, :, t] - expect).max() <= 1e-13 * np.abs(a[:, :, t]).max()
