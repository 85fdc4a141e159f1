"""GPU parity of the separable projection (plg_project_sep.hip): the sums
over x per line, over y per plane and over z per column, in FP64, against
the CPU oracle's per-site DFT (lib/kernels/PLEGMA_kernel_utils.cuh:820-860)
and against the phase-matrix GEMM path (PLG_PROJ_SEP=0).

Cases cover the chunk geometries the host picks: equal and unequal extents
(lines split in 2 or 4 parts, a z extent of 6), momenta whose components
reach L/2 (the residue that is its own negative) and alias (n and n + L), and
the large tables (|n_i| up to 6: seven |n_x| classes, 13 n_y residues).
The standalone FT (plg_ft_apply) runs the same sums on a field read in its
own layout: pinned by the exact-phase DFT of the closed forms on the same
lattices, with partial column groups, IsAccum and an FP32 output.
Tolerances: FP64 1e-12 normwise over momenta; mixed (FP32 contraction) and
FP32 fields 1e-6.
"""
import numpy as np
import pytest

from oracle import oracle as O
from plegma_amd import capi

import gpu_util as G

pytestmark = pytest.mark.gpu

MASK = capi.G_NUCL | capi.G_DELTAPP_ALL
CASES = [  # lattice, |n|^2 bound, source, totalT
    ([8, 8, 8, 16], 16, [2, 5, 7, 3], 12),
    ([8, 4, 6, 8], 9, [1, 3, 5, 6], 8),
    ([16, 16, 8, 8], 36, [9, 2, 6, 1], 6),
]


def _oracle(L, h, src, totalT, q2, rows):
    M = G.oracle_H(O.MESONS16, L, src, totalT, q2, h, h, rows=rows)[:, :1]
    B = G.oracle_H(O.BARYONS, L, src, totalT, q2, h, h, mask=O.G_NUCL | sum(O.G_DELTA32), rows=rows)[:, :1]
    return np.concatenate([M, B], 0)


def _run(view, src, totalT, q2, mode):
    return G.shaped(G.run_correlator(capi.TWOPT, MASK, view, view, src, totalT, q2, mode=mode),
                    capi.TWOPT, MASK, src, totalT, q2)


@pytest.mark.parametrize("case", range(len(CASES)))
def test_separable_projection(case, monkeypatch):
    L, q2, src, totalT = CASES[case]
    G.torch_cuda()
    print(f"\n[gpu] plg_init {L}, |n|^2 <= {q2}", flush=True)
    capi.init(L, 0)
    try:
        t, v = G.make_prop(L, 4242, 3 + case, src[3], 0.5)
        h = G.host(t)
        _, v32 = keep = G.cast_f32(t, v, scaled=True)
        _, _, lt = capi.window(src, totalT)
        rows = sorted({0, lt // 2, lt - 1})
        R = _oracle(L, h, src, totalT, q2, rows)
        print(f"[gpu] FP64 separable, {len(capi.momlist(q2))} momenta", flush=True)
        H64 = _run(v, src, totalT, q2, capi.MODE_F64)
        err = G.rel_err(H64[:, :, rows], R[:, :, rows])
        print(f"[gpu] FP64 separable vs oracle {err:.2e}", flush=True)
        assert err < 1e-12
        print("[gpu] mixed separable", flush=True)
        Hmx = _run(v32, src, totalT, q2, capi.MODE_MIXED)
        err = G.rel_err(Hmx[:, :, rows], R[:, :, rows])
        print(f"[gpu] mixed separable vs oracle {err:.2e}", flush=True)
        assert err < 1e-6
        monkeypatch.setenv("PLG_PROJ_SEP", "0")
        print("[gpu] FP64 phase-matrix GEMM", flush=True)
        Hg = _run(v, src, totalT, q2, capi.MODE_F64)
        err = G.rel_err(H64, Hg)
        print(f"[gpu] separable vs GEMM (all rows) {err:.2e}", flush=True)
        assert err < 1e-12
        del keep
    finally:
        capi.finalize()


def test_separable_is_the_batch_projection():
    """a batch plan on these momenta projects with the separable form, and
    with PLG_PROJ_SEP=0 with the GEMM; both agree"""
    torch = G.torch_cuda()
    L, q2, src, totalT = CASES[0]
    capi.init(L, 0)
    try:
        t, v = G.make_prop(L, 4242, 3, src[3], 0.5)
        moms = np.array(capi.momlist(q2)).reshape(-1, 3).tolist()
        n = capi.correlator_size(capi.TWOPT, MASK, src, totalT, len(moms))
        outs = {}
        for sep in ("1", "0"):
            import os
            os.environ["PLG_PROJ_SEP"] = sep
            try:
                b = capi.Batch(capi.TWOPT, MASK, [v], [(src, totalT, 0, 0)], moms, capi.MODE_F64)
                out = torch.zeros((n, 2), dtype=torch.float64, device="cuda")
                names = [s for s, _ in b.profile(out)]
                print(f"[gpu] PLG_PROJ_SEP={sep}: stages {names}", flush=True)
                assert ("project_sep_f64" in names) == (sep == "1")
                outs[sep] = G.host(out)
                b.close()
            finally:
                del os.environ["PLG_PROJ_SEP"]
        a, g = outs["1"], outs["0"]
        assert np.abs(a - g).max() <= 1e-12 * np.abs(g).max()
    finally:
        capi.finalize()


def test_separable_custom_momentum_list(monkeypatch):
    """a momentum list in no particular order, with repeated momenta and
    momenta aliased by +-L (the same phases): every entry equals the entry of
    its momentum in the createMom result (checked against the oracle above),
    and the GEMM (PLG_PROJ_SEP=0) agrees"""
    torch = G.torch_cuda()
    L, q2, src, totalT = CASES[0]
    capi.init(L, 0)
    try:
        t, v = G.make_prop(L, 4242, 3, src[3], 0.5)
        base = np.array(capi.momlist(q2)).reshape(-1, 3)
        rng = np.random.default_rng(5)
        pick = rng.permutation(len(base))[:40]
        pick = np.concatenate([pick, pick[:5]])            # repeated momenta
        custom = base[pick].copy()
        custom[3] += np.array([L[0], 0, 0])                # aliased by +Lx
        custom[7] -= np.array([0, L[1], L[2]])             # aliased by -Ly, -Lz
        mom_list = custom.tolist()

        def run(moms):
            n = capi.correlator_size(capi.TWOPT, MASK, src, totalT, len(moms))
            out = torch.full((n, 2), float("nan"), dtype=torch.float64, device="cuda")
            capi.correlator(capi.TWOPT, MASK, v, v, src, totalT, moms, capi.MODE_F64, capi.MOMENTUM_SPACE,
                            capi.F64, out)
            h = G.host(out).reshape(n, 2)
            ng, nd, nch = capi.kind_shape(capi.TWOPT, MASK)
            _, _, lt = capi.window(src, totalT)
            return (h[:, 0] + 1j * h[:, 1]).reshape(ng, nd, lt, len(moms), nch)

        ref = run(base.tolist())
        got = run(mom_list)
        err = G.rel_err(got, ref[:, :, :, pick])
        print(f"[gpu] custom list vs createMom entries {err:.2e}", flush=True)
        assert err < 1e-12
        monkeypatch.setenv("PLG_PROJ_SEP", "0")
        gemm = run(mom_list)
        err = G.rel_err(gemm, got)
        print(f"[gpu] custom list, separable vs GEMM {err:.2e}", flush=True)
        assert err < 1e-12
    finally:
        capi.finalize()


def test_static_entry_falls_back_to_the_gemm():
    """plg_correlator_static on a plan whose separable sums would not fit the
    static scratch (64 (n_x, n_y) pairs x 32 z planes per column) but whose
    GEMM partials do: it runs with the GEMM and agrees with plg_correlator
    (separable) to the mixed path's tolerance"""
    torch = G.torch_cuda()
    L, src, totalT = [8, 8, 32, 64], [3, 5, 17, 9], 48
    moms = [[nx, ny, 0] for nx in range(-3, 5) for ny in range(-3, 5)]
    cols = totalT * 80                                   # rows x channels of the two-point kind
    V3 = L[0] * L[1] * L[2]
    sep_bytes = (len(moms) + L[2] * 64) * cols * 16      # sums per momentum + per-plane sums
    c_bytes = cols * V3 * 8                              # FP32 site correlator
    assert sep_bytes + c_bytes > capi.STATIC_SCRATCH_BYTES
    assert c_bytes + 16 * cols * 32 * 32 < capi.STATIC_SCRATCH_BYTES  # GEMM partials, up to 16 K splits
    capi.init(L, 0)
    try:
        t, v = G.make_prop(L, 4242, 9, src[3], 0.5)
        _, v32 = keep = G.cast_f32(t, v, scaled=True)
        n = capi.correlator_size(capi.TWOPT, MASK, src, totalT, len(moms))
        ref = torch.full((n, 2), float("nan"), dtype=torch.float64, device="cuda")
        out = torch.full((n, 2), float("nan"), dtype=torch.float64, device="cuda")
        print("[gpu] plg_correlator (separable) and plg_correlator_static (GEMM)", flush=True)
        capi.correlator(capi.TWOPT, MASK, v32, v32, src, totalT, moms, capi.MODE_MIXED, capi.MOMENTUM_SPACE,
                        capi.F64, ref)
        capi.correlator_static(capi.TWOPT, MASK, v32, v32, src, totalT, moms, capi.MODE_MIXED,
                               capi.MOMENTUM_SPACE, capi.F64, out)
        ng, nd, nch = capi.kind_shape(capi.TWOPT, MASK)
        h, r = G.host(out).reshape(n, 2), G.host(ref).reshape(n, 2)
        got = (h[:, 0] + 1j * h[:, 1]).reshape(ng, nd, totalT, len(moms), nch)
        exp = (r[:, 0] + 1j * r[:, 1]).reshape(ng, nd, totalT, len(moms), nch)
        err = G.rel_err(got, exp)
        print(f"[gpu] static (GEMM) vs separable {err:.2e}", flush=True)
        assert np.isfinite(got).all() and err < 1e-6
        del keep
    finally:
        capi.finalize()


@pytest.mark.parametrize("case", range(len(CASES)))
def test_separable_ft_on_the_field_in_place(case, monkeypatch):
    """plg_ft_apply's separable sums read the field in its own [column][site]
    layout (no copy into tiles): FP64 and FP32 4D fields and a 3D field, on
    the lattices above (unequal extents, components reaching L/2, aliases),
    against the exact-phase DFT of the closed forms and against the GEMM"""
    import closed_forms as cf
    torch = G.torch_cuda()
    L, q2, _, _ = CASES[case]
    V3, T = L[0] * L[1] * L[2], L[3]
    capi.init(L, 0)
    try:
        moms = capi.momlist(q2)
        rng = np.random.default_rng(90 + case)
        for prec, tol, nt, dof in ((capi.F64, 1e-12, T, 3), (capi.F32, 1e-6, T, 5), (capi.F64, 1e-12, 1, 17)):
            f = rng.normal(size=(dof, nt, V3)) + 1j * rng.normal(size=(dof, nt, V3))
            dt = torch.float64 if prec == capi.F64 else torch.float32
            fd = torch.tensor(np.stack([f.real, f.imag], -1), dtype=dt, device="cuda")
            fq = f if prec == capi.F64 else f.astype(np.complex64).astype(np.complex128)
            got = {}
            for sep in ("1", "0"):
                monkeypatch.setenv("PLG_PROJ_SEP", sep)
                out = torch.full((nt, dof, len(moms), 2), float("nan"), dtype=torch.float64, device="cuda")
                print(f"[gpu] L={L} prec={prec} nt={nt} dof={dof} PLG_PROJ_SEP={sep}", flush=True)
                # the separable run passes the list as an int32 array (no conversion)
                ml = np.array(moms, dtype=np.int32) if sep == "1" else moms
                capi.ft_apply(fd, prec, dof, nt, ml, 1, out)
                h = G.host(out).reshape(nt, dof, len(moms), 2)
                got[sep] = h[..., 0] + 1j * h[..., 1]
            mm = -np.array(moms)                                  # exp(+i n.x) = exp(-i (-n).x)
            for t in sorted({0, nt // 2, nt - 1}):
                ref = cf.dft(fq[:, t, :].T, L, [0, 0, 0], mm)     # [mom][dof]
                err = G.rel_err(got["1"][t].T, ref, axis=0)
                print(f"[gpu]   t={t} separable vs DFT {err:.2e}", flush=True)
                assert err < tol
            assert G.rel_err(got["1"], got["0"], axis=2) < tol
            if prec == capi.F64 and nt == T:
                # IsAccum adds to the output; an FP32 output rounds the FP64 sums
                monkeypatch.setenv("PLG_PROJ_SEP", "1")
                g1 = got["1"]
                acc = torch.tensor(np.stack([g1.real, g1.imag], -1), dtype=torch.float64, device="cuda")
                capi.ft_apply(fd, prec, dof, nt, moms, 1, acc, accumulate=True)
                h = G.host(acc).reshape(nt, dof, len(moms), 2)
                assert G.rel_err(h[..., 0] + 1j * h[..., 1], 2 * g1, axis=2) < 1e-14
                o32 = torch.full((nt, dof, len(moms), 2), float("nan"), dtype=torch.float32, device="cuda")
                capi.ft_apply(fd, prec, dof, nt, moms, 1, o32, out_prec=capi.F32)
                h = G.host(o32).reshape(nt, dof, len(moms), 2).astype(np.float64)
                assert G.rel_err(h[..., 0] + 1j * h[..., 1], g1, axis=2) < 1e-6
    finally:
        capi.finalize()
