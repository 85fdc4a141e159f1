"""GPU parity of the 2pt path through the C ABI against the CPU oracle.

Inputs are seeded propagators generated on the device (plg_synth_prop); the
oracle reads the very same bytes copied back to the host.  FP64 path:
normwise relative error <= 1e-12 per (group, dataset, row, channel) over the
momentum list.  Mixed path (FP32 contraction, split-bf16 projection): error
per channel and momentum as stated in each test.
"""
import numpy as np
import pytest

from oracle import oracle as O
from plegma_amd import capi

import gpu_util as G

pytestmark = pytest.mark.gpu

L1 = [8, 8, 8, 16]      # the 8^3x16 lattice of the small two-point workload
Q1 = 4                   # |n|^2 <= 4 : 33 momenta
TOL64 = 1e-12


@pytest.fixture(scope="module", autouse=True)
def ctx():
    G.torch_cuda()
    print("\n[gpu] plg_init", L1, flush=True)
    capi.init(L1, 0)
    yield
    capi.finalize()


@pytest.fixture(scope="module")
def props():
    print("[gpu] synth props", flush=True)
    t1, v1 = G.make_prop(L1, 1234, 1, 0, 0.5)
    t2, v2 = G.make_prop(L1, 1234, 2, 0, 0.5)
    return (t1, v1, G.host(t1)), (t2, v2, G.host(t2))


def test_synth_matches_oracle(props):
    (t1, v1, h1), _ = props
    ref = O.synth_prop(L1, 1234, 1, 0, 0.5)
    np.testing.assert_allclose(h1, ref, rtol=1e-13, atol=1e-15)


def test_momlist_matches_oracle():
    for q in (0, 1, 4, 9, 16):
        assert np.array_equal(np.array(capi.momlist(q)).reshape(-1, 3), O.momlist(q))


@pytest.mark.parametrize("src,totalT", [([0, 0, 0, 0], 16), ([3, 5, 1, 13], 8), ([7, 0, 2, 5], 16)])
def test_window_rules(src, totalT):
    assert capi.window(src, totalT) == O.window(L1, src, totalT)
    _, _, lt = capi.window(src, totalT)
    for j in range(lt):
        assert capi.row_time(src, totalT, j) == O.row_time(L1, src, totalT, j)


@pytest.mark.parametrize("kind,okind,same", [
    (capi.MESONS, O.MESONS, False), (capi.MESONS, O.MESONS, True),
    (capi.MESONS16, O.MESONS16, False), (capi.MESONS_NEW, O.MESONS_NEW, False),
    (capi.MESONS_ALL, O.MESONS_ALL, False)])
def test_mesons_f64(props, kind, okind, same):
    (t1, v1, h1), (t2, v2, h2) = props
    src, totalT = [0, 0, 0, 0], 16
    if same:
        v2, h2 = v1, h1
    print(f"[gpu] mesons kind={kind} same={same}", flush=True)
    H = G.shaped(G.run_correlator(kind, 0, v1, v2, src, totalT, Q1), kind, 0, src, totalT, Q1)
    R = G.oracle_H(okind, L1, src, totalT, Q1, h1, h2)
    assert G.rel_err(H, R) < TOL64


@pytest.mark.parametrize("src,totalT", [([0, 0, 0, 0], 16), ([3, 5, 1, 13], 8)])
@pytest.mark.parametrize("same", [False, True])
def test_baryons_f64(props, src, totalT, same):
    (t1, v1, h1), (t2, v2, h2) = props
    if same:
        v2, h2 = v1, h1
    mask = capi.G_NUCL | capi.G_DELTAPP_ALL
    print(f"[gpu] baryons src={src} totalT={totalT} same={same}", flush=True)
    H = G.shaped(G.run_correlator(capi.BARYONS, mask, v1, v2, src, totalT, Q1),
                 capi.BARYONS, mask, src, totalT, Q1)
    rows = [0, 3, totalT - 1]
    R = G.oracle_H(O.BARYONS, L1, src, totalT, Q1, h1, h2,
                   mask=O.G_NUCL | sum(O.G_DELTA32), rows=rows)
    assert G.rel_err(H[:, :, rows], R[:, :, rows]) < TOL64


def test_baryon_group_subset(props):
    (t1, v1, h1), (t2, v2, h2) = props
    src, totalT = [1, 2, 3, 4], 16
    mask = capi.G_DELTAPP_22
    H = G.shaped(G.run_correlator(capi.BARYONS, mask, v1, v2, src, totalT, 1),
                 capi.BARYONS, mask, src, totalT, 1)
    rows = [0, 11]
    R = G.oracle_H(O.BARYONS, L1, src, totalT, 1, h1, h2, mask=O.G_DELTA32[1], rows=rows)
    assert G.rel_err(H[:, :, rows], R[:, :, rows]) < TOL64


def _twopt_oracle(h, src, totalT, q2, rows):
    M = G.oracle_H(O.MESONS16, L1, src, totalT, q2, h, h, rows=rows)[:, :1]
    B = G.oracle_H(O.BARYONS, L1, src, totalT, q2, h, h, mask=O.G_NUCL | sum(O.G_DELTA32),
                   rows=rows)[:, :1]
    return np.concatenate([M, B], 0)


def test_twopt_f64(props):
    (t1, v1, h1), _ = props
    src, totalT = [2, 1, 0, 9], 12
    mask = capi.G_NUCL | capi.G_DELTAPP_ALL
    H = G.shaped(G.run_correlator(capi.TWOPT, mask, v1, v1, src, totalT, Q1),
                 capi.TWOPT, mask, src, totalT, Q1)
    rows = [0, 6, 11]
    R = _twopt_oracle(h1, src, totalT, Q1, rows)
    assert G.rel_err(H[:, :, rows], R[:, :, rows]) < TOL64


@pytest.mark.parametrize("split,tol", [(capi.SPLIT_BF16X6, 1e-6), (capi.SPLIT_BF16X3, 1e-4)])
def test_twopt_mixed_error(props, split, tol):
    """FP32 contraction + split-bf16 MFMA projection vs the FP64 oracle; error
    per (group, channel, momentum) normalised per time slice."""
    (t1, v1, h1), _ = props
    t32, v32 = G.cast_f32(t1, v1)
    src, totalT = [0, 0, 0, 0], 16
    mask = capi.G_NUCL | capi.G_DELTAPP_ALL
    if split == capi.SPLIT_BF16X6:
        H = G.run_correlator(capi.TWOPT, mask, v32, v32, src, totalT, Q1, mode=capi.MODE_MIXED)
    else:
        moms = capi.momlist(Q1)
        b = capi.Batch(capi.TWOPT, mask, [v32], [(src, totalT, 0, 0)], moms, capi.MODE_MIXED,
                       split=split)
        n, _, _ = b.info()
        torch = G.torch_cuda()
        out = torch.zeros((n, 2), dtype=torch.float64, device="cuda")
        b.run(out)
        h = G.host(out).reshape(n, 2)
        H = h[:, 0] + 1j * h[:, 1]
        b.close()
    H = G.shaped(H, capi.TWOPT, mask, src, totalT, Q1)
    rows = [0, 5, 15]
    R = _twopt_oracle(h1, src, totalT, Q1, rows)
    # normalise by the largest |C| over momenta of each (group, row, channel)
    err = G.rel_err(H[:, :, rows], R[:, :, rows])
    print(f"[gpu] mixed split={split} max normwise rel err {err:.3e}", flush=True)
    assert err < tol


@pytest.mark.parametrize("how", ["cast", "synth"])
def test_mixed_plane_scaling(how):
    """fast decay (lambda = 3) over a wrapped window: the FP32 entries span
    e^-45 and baryons e^-135, below FP32's range; the per-plane exponents keep
    the mixed path within 1e-6 of the FP64 oracle"""
    torch = G.torch_cuda()
    src, totalT, lam = [1, 2, 3, 14], 16, 3.0
    t64, v64 = G.make_prop(L1, 99, 4, src[3], lam)
    h64 = G.host(t64)
    if how == "cast":
        keep, v32 = G.cast_f32(t64, v64, scaled=True)
    else:
        t32 = torch.empty(t64.shape, dtype=torch.float32, device="cuda")
        e = torch.empty(16, dtype=torch.int32, device="cuda")
        v32 = capi.plg_prop(t32.data_ptr(), capi.F32, 0, 16, 0, 16 * 512, e.data_ptr())
        capi.synth_prop(v32, 99, 4, src[3], lam)
        keep = (t32, e)
    mask = capi.G_NUCL | capi.G_DELTAPP_ALL
    H = G.shaped(G.run_correlator(capi.TWOPT, mask, v32, v32, src, totalT, Q1,
                                  mode=capi.MODE_MIXED), capi.TWOPT, mask, src, totalT, Q1)
    rows = [0, 1, 14, 15]          # t = 0, 1 (|t - t0| = 14, 13) and t = 14, 15 (0, 1)
    R = _twopt_oracle(h64, src, totalT, Q1, rows)
    err = G.rel_err(H[:, :, rows], R[:, :, rows])
    print(f"[gpu] plane scaling ({how}) rel err {err:.2e}", flush=True)
    assert err < 1e-6


def test_position_space(props):
    (t1, v1, h1), (t2, v2, h2) = props
    src, totalT = [0, 0, 0, 3], 4
    H = G.run_correlator(capi.MESONS_NEW, 0, v1, v2, src, totalT, 0, space=capi.POSITION_SPACE)
    V3 = 512
    H = H.reshape(1, 1, 4, V3, 10)
    for j in range(4):
        t = O.row_time(L1, src, totalT, j)
        ref = O.site_correlator(O.MESONS_NEW, L1, t, h1, h2)
        assert G.rel_err(H[0, 0, j], ref, axis=0) < 1e-13


def test_batch_matches_single(props):
    (t1, v1, h1), (t2, v2, h2) = props
    torch = G.torch_cuda()
    mask = capi.G_NUCL | capi.G_DELTAPP_ALL
    moms = capi.momlist(Q1)
    jobs_in = [([0, 0, 0, 0], 16, 0), ([3, 5, 1, 13], 8, 1), ([7, 0, 2, 5], 5, 0)]
    jobs, off = [], 0
    for src, totalT, p in jobs_in:
        jobs.append((src, totalT, p, off))
        off += capi.correlator_size(capi.TWOPT, mask, src, totalT, len(moms))
    b = capi.Batch(capi.TWOPT, mask, [v1, v2], jobs, moms, capi.MODE_F64)
    n, nsites, nrows = b.info()
    assert n == off and nrows == 16 + 8 + 5 and nsites == nrows * 512
    out = torch.zeros((n, 2), dtype=torch.float64, device="cuda")
    b.run(out)
    h = G.host(out).reshape(n, 2)
    Hb = h[:, 0] + 1j * h[:, 1]
    for (src, totalT, p), (_, _, _, o) in zip(jobs_in, jobs):
        v = (v1, v2)[p]
        Hs = G.run_correlator(capi.TWOPT, mask, v, v, src, totalT, Q1)
        np.testing.assert_allclose(Hb[o:o + len(Hs)], Hs, rtol=0, atol=1e-12 * np.abs(Hs).max())
    b.close()


def test_batch_graph_replays(props):
    """hipGraph replays interleaved with eager work give the eager result each time."""
    (t1, v1, h1), (t2, v2, h2) = props
    torch = G.torch_cuda()
    t32a, v32a = G.cast_f32(t1, v1)
    t32b, v32b = G.cast_f32(t2, v2)
    mask = capi.G_NUCL | capi.G_DELTAPP_ALL
    moms = capi.momlist(Q1)
    jobs = [([0, 0, 0, 0], 16, 0, 0)]
    jobs.append(([3, 5, 1, 13], 8, 1, capi.correlator_size(capi.TWOPT, mask, [0, 0, 0, 0], 16, len(moms))))
    eager = capi.Batch(capi.TWOPT, mask, [v32a, v32b], jobs, moms, capi.MODE_MIXED)
    graph = capi.Batch(capi.TWOPT, mask, [v32a, v32b], jobs, moms, capi.MODE_MIXED,
                       flags=capi.BATCH_GRAPH)
    n, _, _ = eager.info()
    ref = torch.zeros((n, 2), dtype=torch.float64, device="cuda")
    eager.run(ref)
    ref_h = G.host(ref)
    out = torch.zeros((n, 2), dtype=torch.float64, device="cuda")
    s = torch.cuda.Stream()
    for rep in range(3):
        out.fill_(float("nan"))
        graph.run(out, stream=s)
        s.synchronize()
        assert np.array_equal(G.host(out), ref_h), f"replay {rep}"
        junk = torch.randn(1 << 20, device="cuda").sum()  # eager work between replays
        junk.item()
    eager.close()
    graph.close()


def test_autotuned_batch(props, tmp_path, monkeypatch):
    """the autotuner picks tiles and caches them; results stay within the
    mixed path's tolerance of the untuned plan, and a second plan reads the cache"""
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("PLG_TUNE_CACHE", str(cache))
    (t1, v1, h1), (t2, v2, h2) = props
    torch = G.torch_cuda()
    t32, v32 = G.cast_f32(t1, v1)
    mask = capi.G_NUCL | capi.G_DELTAPP_ALL
    moms = capi.momlist(Q1)
    jobs = [([0, 0, 0, 0], 16, 0, 0)]
    plain = capi.Batch(capi.TWOPT, mask, [v32], jobs, moms, capi.MODE_MIXED)
    tuned = capi.Batch(capi.TWOPT, mask, [v32], jobs, moms, capi.MODE_MIXED,
                       flags=capi.BATCH_TUNE)
    n, _, _ = plain.info()
    a = torch.zeros((n, 2), dtype=torch.float64, device="cuda")
    b = torch.zeros((n, 2), dtype=torch.float64, device="cuda")
    plain.run(a)
    tuned.run(b)
    ha, hb = G.host(a), G.host(b)
    assert np.abs(ha - hb).max() <= 1e-6 * np.abs(ha).max()
    lines = cache.read_text().splitlines()
    assert len(lines) == 1 and " | " in lines[0]
    again = capi.Batch(capi.TWOPT, mask, [v32], jobs, moms, capi.MODE_MIXED,
                       flags=capi.BATCH_TUNE)
    assert len(cache.read_text().splitlines()) == 1   # served from the cache
    for x in (plain, tuned, again):
        x.close()


def test_error_codes(props):
    (t1, v1, h1), _ = props
    torch = G.torch_cuda()
    out = torch.zeros((10, 2), dtype=torch.float64, device="cuda")
    with pytest.raises(capi.PlgError) as e:
        capi.correlator(capi.MESONS, 0, v1, v1, [0, 0, 0, 99], 16, capi.momlist(1),
                        capi.MODE_F64, capi.MOMENTUM_SPACE, capi.F64, out)
    assert e.value.code == -1
    with pytest.raises(capi.PlgError) as e:
        capi.momlist(-1)
    assert e.value.code == -1
    # F64 view handed to the mixed mode
    with pytest.raises(capi.PlgError):
        capi.correlator(capi.MESONS, 0, v1, v1, [0, 0, 0, 0], 16, capi.momlist(1),
                        capi.MODE_MIXED, capi.MOMENTUM_SPACE, capi.F64, out)
    # a window the propagator view does not hold
    tw, vw = G.make_prop(L1, 5, 0, 0, 0.5, t_first=2, nt=4)
    with pytest.raises(capi.PlgError) as e:
        capi.correlator(capi.MESONS, 0, vw, vw, [0, 0, 0, 2], 8, capi.momlist(1),
                        capi.MODE_F64, capi.MOMENTUM_SPACE, capi.F64, out)
    assert e.value.code == -7


def test_window_view(props):
    """a propagator stored only over its source window (t_first, nt) gives the
    same correlator as the full-volume one"""
    (t1, v1, h1), _ = props
    src, totalT = [1, 1, 1, 13], 6     # times 13,14,15,0,1,2
    tw, vw = G.make_prop(L1, 1234, 1, 0, 0.5, t_first=13, nt=6)
    Hf = G.run_correlator(capi.MESONS16, 0, v1, v1, src, totalT, Q1)
    Hw = G.run_correlator(capi.MESONS16, 0, vw, vw, src, totalT, Q1)
    np.testing.assert_allclose(Hw, Hf, rtol=0, atol=1e-13 * np.abs(Hf).max())


# ---- meson-meson (four propagators: light S1, heavy S2)
@pytest.mark.parametrize("same", [False, True])
def test_meson_meson_f64(props, same):
    (t1, v1, h1), (t2, v2, h2) = props
    if same:
        v2, h2 = v1, h1
    src, totalT = [3, 1, 6, 10], 12
    print(f"[gpu] meson-meson f64 same={same}", flush=True)
    H = G.shaped(G.run_correlator(capi.MESON_MESON, 0, v1, v2, src, totalT, Q1),
                 capi.MESON_MESON, 0, src, totalT, Q1)
    rows = [0, 5, 11]
    R = G.oracle_H(O.MESON_MESON, L1, src, totalT, Q1, h1, h2, rows=rows)
    assert G.rel_err(H[:, :, rows], R[:, :, rows]) < TOL64


def test_meson_meson_position_space(props):
    """position space: per-site values of both channel blocks (16 mesons at
    one scale, 4 meson-meson channels at another) land in their slots"""
    (t1, v1, h1), (t2, v2, h2) = props
    src, totalT = [0, 0, 0, 14], 4
    H = G.run_correlator(capi.MESON_MESON, 0, v1, v2, src, totalT, 0, space=capi.POSITION_SPACE)
    V3 = L1[0] * L1[1] * L1[2]
    H = H.reshape(totalT, V3, 20)
    for j in range(totalT):
        t = O.row_time(L1, src, totalT, j)
        ref = O.site_correlator(O.MESON_MESON, L1, t, h1, h2)
        assert G.rel_err(H[j], ref, axis=0) < 1e-12


def test_meson_meson_mixed_two_flavours():
    """light (lambda 0.5) and heavy (lambda 1.2) propagators in FP32 with
    per-plane exponents over a wrapped window: mixed path vs FP64 oracle"""
    torch = G.torch_cuda()
    src, totalT = [2, 3, 1, 12], 16
    tl, vl = G.make_prop(L1, 7, 11, src[3], 0.5)
    th, vh = G.make_prop(L1, 7, 12, src[3], 1.2)
    hl, hh = G.host(tl), G.host(th)
    kl, vl32 = G.cast_f32(tl, vl, scaled=True)
    kh, vh32 = G.cast_f32(th, vh, scaled=True)
    H = G.shaped(G.run_correlator(capi.MESON_MESON, 0, vl32, vh32, src, totalT, Q1,
                                  mode=capi.MODE_MIXED), capi.MESON_MESON, 0, src, totalT, Q1)
    rows = [0, 3, 8, 15]
    R = G.oracle_H(O.MESON_MESON, L1, src, totalT, Q1, hl, hh, rows=rows)
    err = G.rel_err(H[:, :, rows], R[:, :, rows])
    print(f"[gpu] meson-meson mixed rel err {err:.2e}", flush=True)
    assert err < 1e-6


# every projection tile the autotuner may pick, pinned one at a time
PROJ_CODES = [1611, 1521, 1421, 1331, 1332, 1422, 1242, 1322, 1432, 1342, 1222, 1244, 1262, 2322, 2422, 2224, 2324, 2234, 2812, 2424, 2414, 2228]


@pytest.mark.parametrize("code", PROJ_CODES)
def test_projection_tiles(props, code, monkeypatch):
    (t1, v1, h1), _ = props
    monkeypatch.setenv("PLG_PROJ_CODE", str(code))
    keep, v32 = G.cast_f32(t1, v1, scaled=True)
    src, totalT = [5, 3, 1, 7], 16
    mask = capi.G_NUCL | capi.G_DELTAPP_ALL
    print(f"[gpu] projection tile {code}", flush=True)
    H = G.shaped(G.run_correlator(capi.TWOPT, mask, v32, v32, src, totalT, Q1, mode=capi.MODE_MIXED),
                 capi.TWOPT, mask, src, totalT, Q1)
    rows = [0, 9]
    R = _twopt_oracle(h1, src, totalT, Q1, rows)
    assert G.rel_err(H[:, :, rows], R[:, :, rows]) < 1e-6


# ---- standalone projection of a field (PLEGMA_FT::apply)
@pytest.mark.parametrize("prec,tol", [(capi.F64, 1e-12), (capi.F32, 1e-6)])
@pytest.mark.parametrize("sign", [-1, 1])
@pytest.mark.parametrize("sep", ["1", "0"])  # separable sums / phase-matrix GEMM
def test_ft_apply(prec, tol, sign, sep, monkeypatch):
    import closed_forms as cf
    monkeypatch.setenv("PLG_PROJ_SEP", sep)
    torch = G.torch_cuda()
    rng = np.random.default_rng(17)
    dof, T, V3 = 3, L1[3], L1[0] * L1[1] * L1[2]
    f = rng.normal(size=(dof, T, V3)) + 1j * rng.normal(size=(dof, T, V3))
    dt = torch.float64 if prec == capi.F64 else torch.float32
    fd = torch.tensor(np.stack([f.real, f.imag], -1), dtype=dt, device="cuda")
    moms = capi.momlist(Q1)
    out = torch.full((T, dof, len(moms), 2), float("nan"), dtype=torch.float64, device="cuda")
    capi.ft_apply(fd, prec, dof, T, moms, sign, out)
    h = G.host(out).reshape(T, dof, len(moms), 2)
    got = h[..., 0] + 1j * h[..., 1]
    fq = f if prec == capi.F64 else f.astype(np.complex64).astype(np.complex128)
    mm = np.array(moms) * (-1 if sign > 0 else 1)           # exp(+i n.x) = exp(-i (-n).x)
    for t in (0, 7, T - 1):
        ref = cf.dft(fq[:, t, :].T, L1, [0, 0, 0], mm)      # [mom][dof]
        assert G.rel_err(got[t].T, ref, axis=0) < tol


# A momentum list that is not symmetric: duplicates, members without a partner,
# momenta equal to their own negative mod L (zero, components L/2) and one
# outside the Brillouin zone (-9 = -1 mod 8).  The mixed projection pairs
# (n, -n) where it can; every momentum must still project on its own.
IRREGULAR = [(0, 0, 0), (1, 0, 0), (1, 0, 0), (-1, 0, 0), (4, 0, 0), (2, 1, -3), (0, 0, -2),
             (-9, 0, 0), (3, 3, 3), (-3, -3, -3), (4, 4, -4)]


@pytest.mark.parametrize("prec,tol", [(capi.F64, 1e-12), (capi.F32, 1e-6)])
@pytest.mark.parametrize("sep", ["1", "0"])  # separable sums / phase-matrix GEMM
def test_ft_apply_irregular_momenta(prec, tol, sep, monkeypatch):
    import closed_forms as cf
    monkeypatch.setenv("PLG_PROJ_SEP", sep)
    torch = G.torch_cuda()
    rng = np.random.default_rng(29)
    dof, V3 = 2, L1[0] * L1[1] * L1[2]
    f = rng.normal(size=(dof, 1, V3)) + 1j * rng.normal(size=(dof, 1, V3))
    dt = torch.float64 if prec == capi.F64 else torch.float32
    fd = torch.tensor(np.stack([f.real, f.imag], -1), dtype=dt, device="cuda")
    out = torch.full((1, dof, len(IRREGULAR), 2), float("nan"), dtype=torch.float64, device="cuda")
    capi.ft_apply(fd, prec, dof, 1, IRREGULAR, -1, out)
    h = G.host(out).reshape(dof, len(IRREGULAR), 2)
    got = h[..., 0] + 1j * h[..., 1]
    fq = f if prec == capi.F64 else f.astype(np.complex64).astype(np.complex128)
    ref = cf.dft(fq[:, 0, :].T, L1, [0, 0, 0], np.array(IRREGULAR))   # [mom][dof]
    assert G.rel_err(got.T, ref, axis=0) < tol
    # duplicates and aliases give the same numbers
    for a, b in ((1, 2), (3, 7)):
        assert np.abs(got[:, a] - got[:, b]).max() <= 1e-13 * np.abs(got[:, a]).max()


def test_ft_apply_table_cache():
    """the FT keeps the tables of its last momentum list: calls that switch
    between lists, fields and field shapes each match the exact-phase DFT"""
    import closed_forms as cf
    torch = G.torch_cuda()
    rng = np.random.default_rng(31)
    V3 = L1[0] * L1[1] * L1[2]
    lists = [capi.momlist(Q1), IRREGULAR, capi.momlist(Q1)]
    for i, (moms, dof) in enumerate(zip(lists + lists[:1], (3, 2, 3, 3))):  # the last hits the cache
        f = rng.normal(size=(dof, 1, V3)) + 1j * rng.normal(size=(dof, 1, V3))
        fd = torch.tensor(np.stack([f.real, f.imag], -1), dtype=torch.float64, device="cuda")
        out = torch.full((1, dof, len(moms), 2), float("nan"), dtype=torch.float64, device="cuda")
        capi.ft_apply(fd, capi.F64, dof, 1, moms, -1, out)
        h = G.host(out).reshape(dof, len(moms), 2)
        ref = cf.dft(f[:, 0, :].T, L1, [0, 0, 0], np.array(moms))
        assert G.rel_err((h[..., 0] + 1j * h[..., 1]).T, ref, axis=0) < 1e-12, i


def test_mixed_correlator_irregular_momenta(props):
    """Baryon + meson correlator, FP64 and mixed, on the irregular list,
    against the oracle's site correlator projected by the exact-phase DFT of
    the closed forms (the oracle's own DFT takes |n|^2 lists only)."""
    import closed_forms as cf
    (t1, v1, h1), _ = props
    t32, v32 = G.cast_f32(t1, v1)
    src, totalT = [1, 2, 3, 4], 16
    mask = capi.G_NUCL | capi.G_DELTAPP_ALL
    n = capi.correlator_size(capi.TWOPT, mask, src, totalT, len(IRREGULAR))
    torch = G.torch_cuda()
    got = {}
    for mode, v in ((capi.MODE_MIXED, v32), (capi.MODE_F64, v1)):
        out = torch.zeros((n, 2), dtype=torch.float64, device="cuda")
        capi.correlator(capi.TWOPT, mask, v, v, src, totalT, IRREGULAR, mode, capi.MOMENTUM_SPACE, capi.F64, out)
        h = G.host(out).reshape(n, 2)
        got[mode] = (h[:, 0] + 1j * h[:, 1]).reshape(5, -1, len(IRREGULAR), 16)
    _, _, lt = capi.window(src, totalT)
    for j in (0, 5, lt - 1):
        t = capi.row_time(src, totalT, j)
        f = O.site_correlator(O.TWOPT, L1, t, h1, h1, mask=O.G_NUCL | sum(O.G_DELTA32))   # [V3, 5 * 16]
        ref = cf.dft(f, L1, src[:3], np.array(IRREGULAR)).reshape(len(IRREGULAR), 5, 16).transpose(1, 0, 2)
        assert G.rel_err(got[capi.MODE_F64][:, j], ref) < TOL64, j
        assert G.rel_err(got[capi.MODE_MIXED][:, j], ref) < 1e-6, j


def test_concurrent_single_calls(props):
    """plg_correlator from two host threads at once (distinct streams, shared
    workspace, different sizes so the workspace grows under the other call):
    every result equals the one of a call made alone"""
    import threading
    torch = G.torch_cuda()
    (t1, v1, h1), (t2, v2, h2) = props
    moms = capi.momlist(Q1)
    cases = [(capi.TWOPT, capi.G_NUCL | capi.G_DELTAPP_ALL, v1, v1, [0, 0, 0, 0], 16),
             (capi.MESONS16, 0, v1, v2, [3, 5, 1, 13], 8)]
    alone, outs = [], []
    for kind, mask, a, b, src, tt in cases:
        n = capi.correlator_size(kind, mask, src, tt, len(moms))
        o = torch.zeros((n, 2), dtype=torch.float64, device="cuda")
        capi.correlator(kind, mask, a, b, src, tt, moms, capi.MODE_F64, capi.MOMENTUM_SPACE, capi.F64, o)
        alone.append(G.host(o))
        outs.append(torch.zeros_like(o))
    errors = []

    def worker(i):
        kind, mask, a, b, src, tt = cases[i]
        st = torch.cuda.Stream()
        try:
            for _ in range(20):
                outs[i].zero_()
                torch.cuda.synchronize()
                capi.correlator(kind, mask, a, b, src, tt, moms, capi.MODE_F64, capi.MOMENTUM_SPACE, capi.F64,
                                outs[i], st)
                if not np.array_equal(G.host(outs[i]), alone[i]):
                    errors.append(i)
        except Exception as e:  # reported below
            errors.append(repr(e))

    th = [threading.Thread(target=worker, args=(i,)) for i in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
