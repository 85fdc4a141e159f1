"""Benchmark of the two-point correlator path on one MI355X.

Workload (default, "batching"): lattice 32^3 x 64, 16 point sources drawn
from a seed, each with its own time window of 8..32 slices; per source a
seeded propagator stored over its window; 16 meson channels + proton +
Delta++ (Cg1, Cg2, Cg3) = 80 channels; all momenta with |n|^2 <= 16 (257).

Workload "largest" (--workload largest): lattice 64^3 x 128, one source at
the origin, two flavour propagators (light lambda 0.5, heavy lambda 1.2) in
FP32 with per-time-slice exponents (2 x 38.65 GB), 16 heavy-light mesons +
4 meson-meson channels, |n|^2 <= 9 (123 momenta), all 128 time slices; the
FP64 plain path and the error check run on 16 of the time slices
(t = 120..127, 0..7) from FP64 copies of those slices.

A step is one pass over all sources:
  mixed (the metric): FP32 contraction + the projection the autotuner picks
        (the separable FP64 sums over x, y, z, or the split-bf16 (bf16x6)
        phase-matrix GEMM on MFMA), one batched plan replayed as a hipGraph;
  fp64 plain:        FP64 contraction + the same projection form in FP64, one
        plan per source, launched source after source.
The batching workload also times both paths with the projection pinned to the
phase-matrix GEMM (PLG_PROJ_SEP=0) and reports its MFMA use.
value = sites/s of the mixed path (site = (source, x, t) inside the window),
inputs resident in HBM before the timed region.

Also reported: FP64 plain sites/s and the speed-up, the mixed path's error
against the FP64 path per (channel, momentum, time slice), the dominant
kernel against its roofline (HIP events on the launch stream), and the CPU
oracle (reference per-site bodies, FP64, OpenMP) on a bounded sample.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--workload largest|learned]

Workload "udsc" (--workload udsc): lattice 32^3 x 64, one source, four
flavour propagators (u, d, s, c), all 20 flavour products of the udsc baryon
interpolators (200 operators), all 64 time slices, |n|^2 <= 9 (123 momenta);
value = sites/s of the mixed path (FP32 contraction, autotuned projection),
FP64 path and the mixed path's error beside it.

Workload "tetra" (--workload tetra): the four-propagator table engine on
32^3 x 64, one source, light / light / heavy / heavy propagators (as the
reference's "udbb" product), a synthetic table of the reference's shape (16
operators x 768 terms, its four colour patterns in its proportions), 64 time
slices, 123 momenta; FP64 and mixed.

Workload "learned": see run_learned.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "2pt sites/s: mixed-prec batched MFMA vs FP64 plain GPU at ≤1e-6 rel err"
L = [32, 32, 32, 64]
Q2MAX = 16
NSRC = 16
SEED = 20261202
LAMBDA = 0.5
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFS = 2500.0
FP64_PEAK_TFS = 78.6
FP32_VALU_PEAK_TFS = 157.3
# FP32 FMA rate without packed (two-wide) FP32 instructions, which this build
# does not emit
FP32_VALU_UNPACKED_TFS = 78.6

# implemented arithmetic per site of the fused mixed contraction kernel, in
# complex MACs per lane (16 lanes per site), 8 flop per CMAC:
#   meson Gram rows 36; two diquark-class groups (proton + Delta_3,
#   Delta_1 + Delta_2): Q 2 x 36, D 4 x 9, exchange T 2 x 4 x 36, X 4 x 36
CONTRACT_FLOPS_PER_SITE = 8 * 16 * (36 + 2 * 36 + 4 * 9 + 2 * 4 * 36 + 4 * 36)
# meson-meson kernel, complex MACs per lane: meson Gram rows 36, two operator
# pairs x (4 x 27 colour products + 2 x 9 crossed-term products)
MESMES_FLOPS_PER_SITE = 8 * 16 * (36 + 2 * (4 * 27 + 2 * 9))

L4 = [64, 64, 64, 128]
Q2MAX4 = 9
LAMBDA_HEAVY = 1.2


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def sources(rng):
    src = []
    for _ in range(NSRC):
        x = [int(rng.integers(0, L[i])) for i in range(4)]
        w = int(rng.integers(8, 33))
        src.append((x, w))
    return src


def baseline_threads():
    """threads of the CPU baseline: the process's CPU share (OMP_NUM_THREADS
    when set, which the GPU host sets to the cores of one GPU's share, else
    the CPUs this process may run on)"""
    e = os.environ.get("OMP_NUM_THREADS", "")
    if e.isdigit() and int(e) > 0:
        return int(e)
    return max(1, len(os.sched_getaffinity(0)))


def cpu_info():
    """what the CPU baseline ran on: nproc, this process's CPUs, OMP_NUM_THREADS
    and the CPU model (/proc/cpuinfo)"""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "model": model}


def cpu_baseline(P, view, x0, w, rows):
    """Oracle (reference per-site bodies in FP64 + naive per-site DFT, OpenMP)
    on the given rows of the first source, from the host copy P of the source's
    FP64 propagator slices; returns the baseline and the oracle's rows (the
    reference of the error check against the oracle)"""
    from oracle import oracle as O
    nth = baseline_threads()
    mask = O.G_NUCL | sum(O.G_DELTA32)
    V3 = L[0] * L[1] * L[2]
    R, t_used = {}, 0.0
    for j in rows:
        t0 = time.perf_counter()
        H = O.correlator(O.TWOPT, L, x0, w, Q2MAX, P, P, mask=mask, rows=[j], nthreads=nth,
                         view1=view, view2=view)
        t_used += time.perf_counter() - t0
        R[j] = H[:, :, j]
    rate = len(rows) * V3 / t_used
    return {"value": rate, "unit": "sites/s", "cores": nth, "kind": "port", **cpu_info(),
            "sample": f"source 0 of the workload, {len(rows)} time slice(s) x {V3} sites, "
                      f"16 mesons + proton + Delta++ x3, 257 momenta, FP64 reference "
                      f"bodies + per-site DFT, {t_used:.1f} s on {nth} threads"}, R


def cpu_baseline_largest(budget_s=12.0):
    """Oracle meson-meson (FP64 12x12 products per site + per-site DFT,
    OpenMP) on one time slice of the 64^3 x 128 workload"""
    from oracle import oracle as O
    nth = baseline_threads()
    V3 = L4[0] * L4[1] * L4[2]
    t = 3
    PL = O.synth_prop(L4, SEED, 0, 0, LAMBDA, t_first=t, nt=1, nthreads=nth)
    PH = O.synth_prop(L4, SEED, 1, 0, LAMBDA_HEAVY, t_first=t, nt=1, nthreads=nth)
    view = (V3, t, 1)
    t0 = time.perf_counter()
    O.correlator(O.MESON_MESON, L4, [0, 0, 0, 0], L4[3], Q2MAX4, PL, PH, rows=[t], nthreads=nth,
                 view1=view, view2=view)
    used = time.perf_counter() - t0
    return {"value": V3 / used, "unit": "sites/s", "cores": nth, "kind": "port", **cpu_info(),
            "sample": f"time slice t={t} of the 64^3x128 workload ({V3} sites), 16 mesons + "
                      f"4 meson-meson channels, 123 momenta, FP64 12x12 products + per-site "
                      f"DFT, {used:.1f} s"}


def timed(torch, stream, fn, steps):
    """mean milliseconds per call of fn over `steps` calls (HIP events on the
    launch stream) and the host wall time"""
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps, time.perf_counter() - t0


def stage_ms(batch, out, stream, reps=3):
    """median per-stage milliseconds of eager passes of a batch"""
    runs = [batch.profile(out, stream) for _ in range(reps)]
    return {name: float(np.median([dict(r)[name] for r in runs])) for name, _ in runs[0]}


# the sources that make each measured kernel (besides the shared headers and
# the build recipe): counter readings are tied to a hash of these files
KERNEL_SOURCES = {
    "k_baryon": ["csrc/plg_baryon.hip", "csrc/plg_tile.h"],
    "k_mesmes": ["csrc/plg_mesmes.hip", "csrc/plg_tile.h"],
    "k_project_pair": ["csrc/plg_project_bf16.hip"],
    "k_sep_xy": ["csrc/plg_project_sep.hip", "csrc/plg_tile.h"],
    "k_udsc": ["csrc/plg_udsc.hip", "gen_udsc.py"],
    "k_tetra": ["csrc/plg_tetra.hip"],
}
COMMON_SOURCES = ["csrc/plg_internal.h", "csrc/plg_common.h", "../include/plegma_amd.h", "build.py"]


def kernel_family(kernel):
    """'plg::k_baryon<float, ...>' -> 'k_baryon'"""
    name = kernel.split("<")[0].split("(")[0]
    return name.split("::")[-1]


def sources_sha(kernel):
    """hash of the files a kernel is built from (its own sources, the shared
    headers, the build recipe), or None for a kernel without a list"""
    fam = KERNEL_SOURCES.get(kernel_family(kernel))
    if fam is None:
        return None
    h = hashlib.sha256()
    d = os.path.join(ROOT, "plegma_amd")
    for f in sorted(fam) + COMMON_SOURCES:
        h.update(f.encode())
        with open(os.path.normpath(os.path.join(d, f)), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def measured_traffic(workload, stage):
    """HBM bytes per launch of a stage from the committed counter summary
    (profiles/traffic.json): (bytes, kernel, note), bytes None when the
    counters were taken on other sources of that kernel than these"""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None, None, "no counter summary"
    e = d.get(workload, {}).get(stage)
    if not e:
        return None, None, "stage not in the counter summary"
    now = sources_sha(e["kernel"])
    if now is None or e.get("sources_sha") != now:
        return None, e.get("kernel"), (f"stale: counters of {e['kernel']} sources {e.get('sources_sha')}, "
                                       f"this build is {now}")
    return e["bytes"], e["kernel"], f"rocprofv3 FETCH_SIZE/WRITE_SIZE of {e['kernel']} " \
                                    f"(profiles/traffic.json, kernel sources {now})"


def roof(kernel, t_s, nsites, flops_site, bytes_site, bound_compute, cpeak, workload, algorithmic):
    """a kernel against its roofline.  The binding roof is the one the
    kernel's arithmetic intensity (flops per algorithmic HBM byte) puts it
    under: compute when the intensity exceeds the ridge peak/HBM, else HBM;
    the other roof is reported beside it."""
    tf = nsites * flops_site / t_s / 1e12
    gbs = nsites * bytes_site / t_s / 1e9
    ai = flops_site / bytes_site
    ridge = cpeak * 1e12 / (HBM_PEAK_GBS * 1e9)
    comp = {"achieved": tf, "peak": cpeak, "unit": "TFLOP/s", "frac": tf / cpeak, "flops_per_site": flops_site}
    if bound_compute == "valu":
        comp["peak_unpacked_fp32"] = FP32_VALU_UNPACKED_TFS
        comp["frac_unpacked_fp32"] = tf / FP32_VALU_UNPACKED_TFS
    hbm = {"achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
           "bytes_per_site": bytes_site}
    tr, kname, note = measured_traffic(workload, kernel)
    if ai > ridge:
        r = {"kernel": kernel, "bound": bound_compute, **comp, "hbm": hbm}
    elif bound_compute == "hbm":  # a kernel with no arithmetic to price
        r = {"kernel": kernel, "bound": "hbm", **hbm}
    else:
        r = {"kernel": kernel, "bound": "hbm", **hbm, bound_compute: comp}
    r.update({"traffic": tr, "traffic_source": note,
              "arithmetic_intensity": ai, "ridge": ridge, "algorithmic": algorithmic,
              "time_ms": t_s * 1e3})
    if kname:
        r["kernel_symbol"] = kname
    return r


def proj_dtype(st_ms, split):
    """the arithmetic of a mixed plan, from its stage names (the autotuner
    picks the separable FP64 projection or the split-bf16 GEMM)"""
    if any(k.startswith("project_sep") for k in st_ms):
        return "f32 contraction + fp64 separable projection (fp64 sums)"
    return f"f32 contraction + bf16x{split} MFMA projection (fp32 acc, fp64 sums)"


def roofline(st_ms, nsites, nmom, nch, split, contract_flops, prop_bytes, workload):
    """the dominant stage against its roofline: a contraction (FP32 VALU; HBM:
    compulsory propagator read + site-correlator write) or the projection
    (BF16 MFMA, useful flops; HBM: site-correlator read)"""
    dom = max(st_ms, key=st_ms.get)
    t = st_ms[dom] / 1e3
    if dom.startswith("project_sep"):
        return roof(dom, t, nsites, 0, nch * 8, "hbm", 1.0, workload,
                    f"{nch * 8} B FP32 site correlator read per site")
    if dom.startswith("contract"):
        return roof(dom, t, nsites, contract_flops, prop_bytes + nch * 8, "valu", FP32_VALU_PEAK_TFS, workload,
                    f"{contract_flops} implemented FP32 flop; {prop_bytes} B FP32 propagator read + "
                    f"{nch * 8} B site correlator write per site")
    r = roof(dom, t, nsites, 8 * nmom * nch, nch * 8, "mfma", BF16_PEAK_TFS, workload,
             f"useful 8 flop x {nmom} momenta x {nch} channels per site; {nch * 8} B site correlator read")
    r["executed_frac"] = r.get("frac", 0) * split
    return r


def rel_err_block(A, B):
    """[group, momentum] max over time slices and channels of
    |A - B| / max_momenta |B|; A, B: [group, row, momentum, channel]"""
    den = np.abs(B).max(axis=2, keepdims=True)
    den = np.where(den == 0, 1.0, den)
    return (np.abs(A - B) / den).max(axis=(1, 3))


def shell_pools(q2, minsize=24):
    """|n|^2 shells pooled in increasing order until each pool holds at least
    `minsize` momenta (a remainder joins the last pool): boolean masks"""
    q2 = np.asarray(q2)
    pools, cur = [], np.zeros(len(q2), bool)
    for q in np.unique(q2):
        cur = cur | (q2 == q)
        if cur.sum() >= minsize:
            pools.append(cur)
            cur = np.zeros(len(q2), bool)
    if cur.any():
        if pools:
            pools[-1] = pools[-1] | cur
        else:
            pools.append(cur)
    return pools


def shell_err_block(A, B, moms):
    """[group, momentum] max over time slices and channels of |A - B| / rms
    of |B| over the momenta of similar |n| (|n|^2 shells pooled until a pool
    holds at least 24 momenta): the error per channel and momentum at the
    scale of a correlator at that |n| (the n = 0 value does not set the scale
    at |n| > 0; an accidental zero at one momentum, or a shell of one or six
    momenta, does not either); A, B: [group, row, momentum, channel]"""
    q2 = (np.asarray(moms).reshape(-1, 3) ** 2).sum(1)
    den = np.empty(B.shape)
    for sel in shell_pools(q2):
        den[:, :, sel, :] = np.sqrt((np.abs(B[:, :, sel, :]) ** 2).mean(axis=2, keepdims=True))
    den = np.where(den == 0, 1.0, den)
    return (np.abs(A - B) / den).max(axis=(1, 3))


def point_err(A, B):
    """pointwise |A - B| / |B| over nonzero B: median, 99.9th percentile, max"""
    b = np.abs(B).reshape(-1)
    q = np.abs(A - B).reshape(-1)[b > 0] / b[b > 0]
    return {"median": float(np.median(q)), "p99.9": float(np.quantile(q, 0.999)), "max": float(q.max())}


def err_summary(A, B, moms):
    """normwise, per (channel, momentum) at the shell scale, and pointwise
    errors of A against the reference B ([group, row, momentum, channel])"""
    E = rel_err_block(A, B)
    S = shell_err_block(A, B, moms)
    return {"normwise_max": float(E.max()), "normwise_median": float(np.median(E)),
            "per_channel_momentum_max": float(S.max()),
            "per_channel_momentum_median": float(np.median(S)),
            "pointwise": point_err(A, B)}


ERR_DEFS = {"normwise": "|C - C_ref| / max over momenta |C_ref| per (group, row, channel)",
            "per_channel_momentum": "|C - C_ref| per (group, row, channel, momentum) / rms of |C_ref| "
                                    "over the momenta of similar |n| (|n|^2 shells pooled to >= 24 momenta)",
            "pointwise": "|C - C_ref| / |C_ref| per element (unbounded at accidental zeros of C_ref)"}


def run_gemm_projection(args, torch, capi, stream, srcs, v32, v64, jobs, sizes, moms, kind, mask, dev,
                        nsites, nout, out_ref):
    """the same batch with the projection pinned to the phase-matrix GEMM
    (PLG_PROJ_SEP=0: split-bf16 MFMA for the mixed path, FP64 MFMA for the
    plain path): its step time, its MFMA use and the like-for-like ratio"""
    os.environ["PLG_PROJ_SEP"] = "0"
    try:
        mixed = capi.Batch(kind, mask, v32, jobs, moms, capi.MODE_MIXED, split=args.split,
                           flags=capi.BATCH_GRAPH | capi.BATCH_TUNE)
        out = torch.zeros((nout, 2), dtype=torch.float64, device=dev)
        mixed.run(out, stream)
        ms_mx, _ = timed(torch, stream, lambda: mixed.run(out, stream), args.steps)
        st = stage_ms(mixed, out, stream)
        d = (out - out_ref).abs().max().item() / max(out_ref.abs().max().item(), 1e-300)
        mixed.close()
        plain = [capi.Batch(kind, mask, [v64[s]], [(x0, w, 0, 0)], moms, capi.MODE_F64, flags=capi.BATCH_TUNE)
                 for s, (x0, w) in enumerate(srcs)]
        outs = [torch.zeros((n, 2), dtype=torch.float64, device=dev) for n in sizes]

        def run_plain():
            for b, o in zip(plain, outs):
                b.run(o, stream)

        run_plain()
        k64 = args.fp64_steps or max(1, args.steps // 2)
        ms_64, _ = timed(torch, stream, run_plain, k64)
        for b in plain:
            b.close()
    finally:
        del os.environ["PLG_PROJ_SEP"]
    pname = next((k for k in st if k.startswith("project")), None)
    r = {"note": "PLG_PROJ_SEP=0: the projection as the phase-matrix GEMM on MFMA",
         "mixed": {"value": nsites / (ms_mx / 1e3), "ms_per_step": ms_mx, "stages_ms": st,
                   "dtype": proj_dtype(st, args.split),
                   "max_abs_diff_vs_autotuned_rel": d},
         "fp64_plain": {"value": nsites / (ms_64 / 1e3), "ms_per_step": ms_64, "steps": k64},
         "speedup_vs_fp64": ms_64 / ms_mx}
    if pname:
        t = st[pname] / 1e3
        r["projection_mfma"] = roof(pname, t, nsites, 8 * len(moms) * 80, 80 * 8, "mfma", BF16_PEAK_TFS,
                                    "batching", f"useful 8 flop x {len(moms)} momenta x 80 channels per site; "
                                                f"{80 * 8} B site correlator read")
        pm = r["projection_mfma"]
        comp = pm if pm["bound"] == "mfma" else pm["mfma"]
        pm["mfma_executed_frac"] = comp["frac"] * args.split  # split products x useful flops
    return r


def run_batching(args, torch, capi, dev, stream):
    capi.init(L, 0)
    V3 = L[0] * L[1] * L[2]
    rng = np.random.default_rng(SEED)
    srcs = sources(rng)
    moms = capi.momlist(Q2MAX)
    mask = capi.G_NUCL | capi.G_DELTAPP_ALL
    kind = capi.TWOPT

    log("synthesising propagators")
    keep, v64, v32 = [], [], []
    for s, (x0, w) in enumerate(srcs):
        t64 = torch.empty((144, w * V3, 2), dtype=torch.float64, device=dev)
        a = capi.plg_prop(t64.data_ptr(), capi.F64, x0[3], w, 0, w * V3)
        capi.synth_prop(a, SEED, s, x0[3], LAMBDA)
        # FP32 copy with per-time-slice power-of-two exponents (exact rescaling)
        t32 = torch.empty((144, w * V3, 2), dtype=torch.float32, device=dev)
        e32 = torch.empty(w, dtype=torch.int32, device=dev)
        b = capi.plg_prop(t32.data_ptr(), capi.F32, x0[3], w, 0, w * V3, e32.data_ptr())
        capi.cast_prop(b, a)
        keep += [t64, t32, e32]
        v64.append(a)
        v32.append(b)
    torch.cuda.synchronize()

    jobs, off, sizes = [], 0, []
    for s, (x0, w) in enumerate(srcs):
        n = capi.correlator_size(kind, mask, x0, w, len(moms))
        jobs.append((x0, w, s, off))
        sizes.append(n)
        off += n
    log("planning + autotuning the mixed batch")
    mixed = capi.Batch(kind, mask, v32, jobs, moms, capi.MODE_MIXED, split=args.split,
                       flags=capi.BATCH_GRAPH | capi.BATCH_TUNE)
    nout, nsites, nrows = mixed.info()
    out_mx = torch.zeros((nout, 2), dtype=torch.float64, device=dev)

    log(f"mixed: {nrows} rows, {nsites} sites; warmup {args.warmup}")
    for _ in range(args.warmup):
        mixed.run(out_mx, stream)
    ms_mx, _ = timed(torch, stream, lambda: mixed.run(out_mx, stream), args.steps)
    log(f"mixed: {ms_mx:.3f} ms/step")
    st_ms = stage_ms(mixed, out_mx, stream)

    res = {"metric": METRIC, "value": nsites / (ms_mx / 1e3), "unit": "sites/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_mx,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": proj_dtype(st_ms, args.split),
           "data": "synthetic",
           "config": {"workload": "batching: lattice 32^3x64, 16 seeded sources, windows 8-32, "
                                  "16 mesons + proton + Delta++ (3x16), |n|^2<=16 (257 momenta)",
                      "lattice": L, "sources": NSRC, "windows": [w for _, w in srcs],
                      "sites": nsites, "channels": 80, "nmom": len(moms),
                      "parallelism": "single GPU"}}

    # FP64 plain path: one plan per source, launched one after another
    if not args.no_fp64:
        plain = [capi.Batch(kind, mask, [v64[s]], [(x0, w, 0, 0)], moms, capi.MODE_F64,
                            flags=capi.BATCH_TUNE)
                 for s, (x0, w) in enumerate(srcs)]
        outs64 = [torch.zeros((n, 2), dtype=torch.float64, device=dev) for n in sizes]

        def run_plain():
            for b, o in zip(plain, outs64):
                b.run(o, stream)

        run_plain()
        k64 = args.fp64_steps or max(1, args.steps // 2)
        ms_64, _ = timed(torch, stream, run_plain, k64)
        log(f"fp64 plain: {ms_64:.3f} ms/step")
        st64 = {name: ms for name, ms in plain[0].profile(outs64[0], stream)}
        torch.cuda.synchronize()
        hm = out_mx.cpu().numpy()
        hm = hm[:, 0] + 1j * hm[:, 1]
        As, Bs = [], []
        for (x0, w, s, o), n in zip(jobs, sizes):
            b = outs64[s].cpu().numpy()
            As.append(hm[o:o + n].reshape(5, w, len(moms), 16))
            Bs.append((b[:, 0] + 1j * b[:, 1]).reshape(5, w, len(moms), 16))
        A = np.concatenate(As, axis=1)
        B = np.concatenate(Bs, axis=1)
        res["fp64_plain"] = {"value": nsites / (ms_64 / 1e3), "unit": "sites/s",
                             "ms_per_step": ms_64, "steps": k64,
                             "stages_ms_source0": st64, "sites_source0": srcs[0][1] * V3,
                             "algorithm": "the mixed path's: factorised contraction, the autotuner's "
                                          "projection (separable sums or the (n, -n)-paired GEMM); "
                                          "FP64 throughout",
                             "projection": "separable" if "project_sep_f64" in st64 else "gemm"}
        res["speedup_vs_fp64"] = ms_64 / ms_mx
        res["rel_err_vs_fp64"] = {**err_summary(A, B, moms),
                                  "per_group_max": [float(x) for x in rel_err_block(A, B).max(axis=1)],
                                  "scope": "every source, time slice, group, channel and momentum",
                                  "definitions": ERR_DEFS}
        del plain, A, B, As, Bs
        res["gemm_projection"] = run_gemm_projection(args, torch, capi, stream, srcs, v32, v64, jobs, sizes,
                                                     moms, kind, mask, dev, nsites, nout, out_mx)
        res["dropin"] = run_dropin(args, torch, capi, stream, srcs, v32, v64, sizes, moms, kind, mask, dev)

    res["stages_ms"] = st_ms
    if not args.no_cpu:
        # the oracle on rows of source 0, from the same FP64 propagator bytes
        x0, w = srcs[0]
        rows = list(range(min(w, 6)))
        planes = sorted({(capi.row_time(x0, w, j) - x0[3]) % L[3] for j in rows})
        assert planes == list(range(len(planes))), planes
        np_ = len(planes)
        P = np.ascontiguousarray(keep[0].view(144, w, V3, 2)[:, :np_].cpu().numpy().reshape(-1))
        log("cpu baseline (oracle on source 0)")
        res["cpu_baseline"], R = cpu_baseline(P, (np_ * V3, x0[3], np_), x0, w, rows)
        Ro = np.stack([R[j] for j in rows], axis=2)[:, 0]            # [group, row, mom, ch]
        hmx = out_mx.cpu().numpy()
        hmx = (hmx[:, 0] + 1j * hmx[:, 1])[:sizes[0]].reshape(5, w, len(moms), 16)[:, rows]
        err = {"rows": rows, "source": 0, "mixed": err_summary(hmx, Ro, moms), "definitions": ERR_DEFS}
        if not args.no_fp64:
            h64 = outs64[0].cpu().numpy()
            h64 = (h64[:, 0] + 1j * h64[:, 1]).reshape(5, w, len(moms), 16)[:, rows]
            err["fp64_normwise_max"] = float(rel_err_block(h64, Ro).max())
        res["rel_err_vs_oracle"] = err
    res["roofline"] = roofline(st_ms, nsites, len(moms), 80, args.split, CONTRACT_FLOPS_PER_SITE,
                               144 * 8, "batching")
    mixed.close()
    del keep, v64, v32
    if not args.no_cpu:
        try:
            res["config_8x8x8x16"] = run_small(torch, capi, dev, stream)
        except Exception as e:  # the headline line is printed regardless
            res["config_8x8x8x16"] = {"error": repr(e)}
    return res


def run_small(torch, capi, dev, stream, reps=20):
    """The reference's CPU-runnable case: lattice 8^3x16, one point source at
    the origin, 16 mesons + the proton (16 spin projections), |n|^2 <= 4 (33
    momenta): the oracle (FP64 reference bodies + per-site DFT) on all 16 time
    slices and the GPU paths through plg_correlator (the drop-in call), with
    the errors of both GPU paths against the oracle"""
    from oracle import oracle as O
    Ls = [8, 8, 8, 16]
    capi.init(Ls, 0)
    V3, T = Ls[0] * Ls[1] * Ls[2], Ls[3]
    src = [0, 0, 0, 0]
    moms = capi.momlist(4)
    kind, mask = capi.TWOPT, capi.G_NUCL
    t64 = torch.empty((144, T * V3, 2), dtype=torch.float64, device=dev)
    v64 = capi.plg_prop(t64.data_ptr(), capi.F64, 0, T, 0, T * V3)
    capi.synth_prop(v64, SEED, 0, 0, LAMBDA)
    t32 = torch.empty((144, T * V3, 2), dtype=torch.float32, device=dev)
    e32 = torch.empty(T, dtype=torch.int32, device=dev)
    v32 = capi.plg_prop(t32.data_ptr(), capi.F32, 0, T, 0, T * V3, e32.data_ptr())
    capi.cast_prop(v32, v64)
    n = capi.correlator_size(kind, mask, src, T, len(moms))
    out = {}
    res = {"workload": "lattice 8^3x16, source at the origin, 16 mesons + proton, |n|^2<=4 (33 momenta)",
           "sites": V3 * T}
    for name, v, mode in (("fp64", v64, capi.MODE_F64), ("mixed", v32, capi.MODE_MIXED)):
        o = torch.empty((n, 2), dtype=torch.float64, device=dev)
        capi.correlator(kind, mask, v, v, src, T, moms, mode, capi.MOMENTUM_SPACE, capi.F64, o, stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            capi.correlator(kind, mask, v, v, src, T, moms, mode, capi.MOMENTUM_SPACE, capi.F64, o, stream)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / reps
        h = o.cpu().numpy()
        out[name] = (h[:, 0] + 1j * h[:, 1]).reshape(2, T, len(moms), 16)
        res[f"gpu_{name}"] = {"value": V3 * T / (ms / 1e3), "unit": "sites/s", "ms_per_call": ms,
                              "note": "plg_correlator, one call per correlator (synchronising)"}
    P = np.ascontiguousarray(t64.cpu().numpy().reshape(-1))
    nth = baseline_threads()
    runs, t_cpu = 0, 0.0
    while t_cpu < 1.0:  # repeated to about a second of CPU time
        t0 = time.perf_counter()
        R = O.correlator(O.TWOPT, Ls, src, T, 4, P, P, mask=O.G_NUCL, nthreads=nth)
        t_cpu += time.perf_counter() - t0
        runs += 1
    R = R[:, 0]
    res["cpu_baseline"] = {"value": runs * V3 * T / t_cpu, "unit": "sites/s", "cores": nth, "kind": "port",
                           "sample": f"the whole workload ({V3 * T} sites) {runs} times, {t_cpu:.2f} s"}
    for name in ("fp64", "mixed"):
        A = out[name][:, :, :, :]
        res[f"rel_err_{name}_vs_oracle"] = err_summary(A, R, moms)
    capi.finalize()
    return res


def run_dropin(args, torch, capi, stream, srcs, v32, v64, sizes, moms, kind, mask, dev):
    """the path a PLEGMA driver takes: one plg_correlator call per source (the
    C ABI under PLEGMA_Correlator::contract*), as plegma/Calc_2pt.cpp:121-132
    loops over sources; each call synchronises its stream, plans are cached"""
    V3 = L[0] * L[1] * L[2]
    outs = [torch.empty((n, 2), dtype=torch.float64, device=dev) for n in sizes]
    nsites = sum(w for _, w in srcs) * V3
    res = {"note": "plg_correlator source by source (cached plans, autotuned tiles, one stream "
                   "synchronisation per call)"}
    for name, views, mode in (("mixed", v32, capi.MODE_MIXED), ("fp64", v64, capi.MODE_F64)):
        def step():
            for (x0, w), v, o in zip(srcs, views, outs):
                capi.correlator(kind, mask, v, v, x0, w, moms, mode, capi.MOMENTUM_SPACE, capi.F64, o, stream)
        step()  # plans and tuning of every source
        k = max(1, args.steps // 2)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(k):
            step()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / k
        log(f"drop-in {name}: {ms:.3f} ms/step")
        res[name] = {"value": nsites / (ms / 1e3), "unit": "sites/s", "ms_per_step": ms, "steps": k}
    return res


def run_largest(args, torch, capi, dev, stream):
    capi.init(L4, 0)
    V3 = L4[0] * L4[1] * L4[2]
    T = L4[3]
    src = [0, 0, 0, 0]
    moms = capi.momlist(Q2MAX4)
    kind = capi.MESON_MESON
    log("synthesising two FP32 flavour propagators (2 x 38.65 GB)")
    keep, v32 = [], []
    for fl, lam in enumerate((LAMBDA, LAMBDA_HEAVY)):
        t32 = torch.empty((144, T * V3, 2), dtype=torch.float32, device=dev)
        e32 = torch.empty(T, dtype=torch.int32, device=dev)
        v = capi.plg_prop(t32.data_ptr(), capi.F32, 0, T, 0, T * V3, e32.data_ptr())
        capi.synth_prop(v, SEED, fl, 0, lam)
        keep += [t32, e32]
        v32.append(v)
    torch.cuda.synchronize()
    nout = capi.correlator_size(kind, 0, src, T, len(moms))
    log("planning + autotuning the mixed plan")
    mixed = capi.Batch(kind, 0, v32, [(src, T, 0, 0, 1)], moms, capi.MODE_MIXED, split=args.split,
                       flags=capi.BATCH_GRAPH | capi.BATCH_TUNE)
    _, nsites, nrows = mixed.info()
    out_mx = torch.zeros((nout, 2), dtype=torch.float64, device=dev)
    log(f"mixed: {nrows} rows, {nsites} sites; warmup {args.warmup}")
    for _ in range(args.warmup):
        mixed.run(out_mx, stream)
    ms_mx, _ = timed(torch, stream, lambda: mixed.run(out_mx, stream), args.steps)
    log(f"mixed: {ms_mx:.3f} ms/step")
    st_ms = stage_ms(mixed, out_mx, stream)
    res = {"metric": METRIC, "value": nsites / (ms_mx / 1e3), "unit": "sites/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_mx,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": proj_dtype(st_ms, args.split),
           "data": "synthetic",
           "config": {"workload": "largest: lattice 64^3x128, one source, light + heavy FP32 "
                                  "propagators, 16 heavy-light mesons + 4 meson-meson, "
                                  "|n|^2<=9 (123 momenta)",
                      "lattice": L4, "sources": 1, "sites": nsites, "channels": 20,
                      "nmom": len(moms), "parallelism": "single GPU"}}
    mixed.close()

    if not args.no_fp64:
        # FP64 copies of 16 time slices t = 120..127, 0..7 (a view that wraps),
        # contracted as the window of a source at t0 = 120 (rows j -> t by the
        # window rule: t = 0..7 first, then 120..127)
        nt, tf = 16, 120
        v64, k64 = [], []
        for fl, lam in enumerate((LAMBDA, LAMBDA_HEAVY)):
            t64 = torch.empty((144, nt * V3, 2), dtype=torch.float64, device=dev)
            v = capi.plg_prop(t64.data_ptr(), capi.F64, tf, nt, 0, nt * V3)
            capi.synth_prop(v, SEED, fl, 0, lam)
            k64.append(t64)
            v64.append(v)
        src64 = [0, 0, 0, tf]
        n64 = capi.correlator_size(kind, 0, src64, nt, len(moms))
        plain = capi.Batch(kind, 0, v64, [(src64, nt, 0, 0, 1)], moms, capi.MODE_F64,
                           flags=capi.BATCH_TUNE)
        out64 = torch.zeros((n64, 2), dtype=torch.float64, device=dev)
        plain.run(out64, stream)
        ms_64, _ = timed(torch, stream, lambda: plain.run(out64, stream), max(1, args.steps // 2))
        st64 = stage_ms(plain, out64, stream, reps=1)
        torch.cuda.synchronize()
        a = out_mx.cpu().numpy()
        b = out64.cpu().numpy()
        A = (a[:, 0] + 1j * a[:, 1]).reshape(1, T, len(moms), 20)
        B = (b[:, 0] + 1j * b[:, 1]).reshape(1, nt, len(moms), 20)
        rows = [capi.row_time(src64, nt, j) for j in range(nt)]
        E = np.concatenate([rel_err_block(A[:, rows, :, :16], B[..., :16]),
                            rel_err_block(A[:, rows, :, 16:], B[..., 16:])], 0)
        sites64 = nt * V3
        res["fp64_plain"] = {"value": sites64 / (ms_64 / 1e3), "unit": "sites/s",
                             "ms_per_step": ms_64, "sites": sites64, "stages_ms": st64,
                             "sample": "time slices 120..127, 0..7"}
        res["speedup_vs_fp64"] = res["value"] / res["fp64_plain"]["value"]
        res["rel_err_vs_fp64"] = {"max": float(E.max()), "median": float(np.median(E)),
                                  "per_group_max": {"mesons": float(E[0].max()),
                                                    "meson_meson": float(E[1].max())},
                                  "definition": "max over the 16 compared time slices and "
                                                "channels of |C_mixed - C_fp64| / "
                                                "max_momenta |C_fp64|, per channel block and "
                                                "momentum"}
        plain.close()
        del k64
    res["roofline"] = roofline(st_ms, nsites, len(moms), 20, args.split, MESMES_FLOPS_PER_SITE,
                               2 * 144 * 8, "largest")
    res["stages_ms"] = st_ms
    if not args.no_cpu:
        log("cpu baseline")
        res["cpu_baseline"] = cpu_baseline_largest()
    return res


UDSC_SRC = [3, 5, 7, 10]
UDSC_LAMBDAS = (0.5, 0.55, 0.8, 1.2)


def udsc_work_per_site():
    """implemented work of the udsc kernel per site, summed over the 20
    products: complex MACs (36 per (m0m1, m2m3) quad for the colour diquark,
    9 per distinct term, 1 per operator entry) and the compulsory bytes
    (three propagators read per product)"""
    from plegma_amd import gen_udsc
    cmac = 0
    nops = 0
    for _, ops in gen_udsc.products():
        terms = set(m for _, t in ops for m in t)
        quads = set(m[:4] for m in terms)
        cmac += 36 * len(quads) + 9 * len(terms) + sum(len(t) for _, t in ops)
        nops += len(ops)
    return cmac, nops


def cpu_baseline_udsc(budget_s=12.0):
    """Oracle udsc (the reference's per-term epsilon sums in FP64 + per-site
    DFT, OpenMP) on an 8^3 spatial lattice; the per-site work does not depend
    on the lattice size"""
    from oracle import oracle as O
    from plegma_amd import gen_udsc
    nth = baseline_threads()
    Ls = [8, 8, 8, 16]
    V3 = Ls[0] * Ls[1] * Ls[2]
    props = {f: O.synth_prop(Ls, SEED, i, 0, lam, nthreads=nth) for i, (f, lam) in enumerate(zip("udsc", UDSC_LAMBDAS))}
    tables = gen_udsc.products()
    done, used = 0, 0.0
    for totalT in (1, 2, 4, 8, 16):
        t0 = time.perf_counter()
        O.udsc_correlator(Ls, [1, 2, 3, 0], totalT, Q2MAX4, props, tables, nthreads=nth)
        used = time.perf_counter() - t0
        done = totalT
        if used * 2.2 > budget_s:
            break
    return {"value": done * V3 / used, "unit": "sites/s", "cores": nth, "kind": "port", **cpu_info(),
            "sample": f"{done} time slice(s) x {V3} sites of an 8^3x16 lattice, 20 flavour products "
                      f"(200 operators), 123 momenta, FP64 per-term epsilon sums + per-site DFT, {used:.1f} s"}


def run_udsc(args, torch, capi, dev, stream):
    capi.init(L, 0)
    V3 = L[0] * L[1] * L[2]
    T = L[3]
    moms = capi.momlist(Q2MAX4)
    log("synthesising four FP64 flavour propagators and their FP32 copies")
    v64, v32, keep = [], [], []
    for fl, lam in enumerate(UDSC_LAMBDAS):
        t64 = torch.empty((144, T * V3, 2), dtype=torch.float64, device=dev)
        v = capi.plg_prop(t64.data_ptr(), capi.F64, 0, T, 0, T * V3)
        capi.synth_prop(v, SEED, 40 + fl, UDSC_SRC[3], lam)
        t32 = torch.empty((144, T * V3, 2), dtype=torch.float32, device=dev)
        e32 = torch.empty(T, dtype=torch.int32, device=dev)
        w = capi.plg_prop(t32.data_ptr(), capi.F32, 0, T, 0, T * V3, e32.data_ptr())
        capi.cast_prop(w, v)
        keep += [t64, t32, e32]
        v64.append(v)
        v32.append(w)
    torch.cuda.synchronize()
    n = capi.baryons_udsc_size(*v32, UDSC_SRC, T, len(moms))
    out_mx = torch.zeros((n, 2), dtype=torch.float64, device=dev)
    out64 = torch.zeros((n, 2), dtype=torch.float64, device=dev)

    def mixed():
        capi.baryons_udsc(*v32, UDSC_SRC, T, moms, capi.MODE_MIXED, capi.MOMENTUM_SPACE, capi.F64, out_mx,
                          stream=stream)

    def plain():
        capi.baryons_udsc(*v64, UDSC_SRC, T, moms, capi.MODE_F64, capi.MOMENTUM_SPACE, capi.F64, out64,
                          stream=stream)

    nsites = V3 * T
    log(f"mixed: {nsites} sites x 20 products; warmup {args.warmup}")
    for _ in range(args.warmup):
        mixed()
    ms_mx, _ = timed(torch, stream, mixed, args.steps)
    log(f"mixed: {ms_mx:.3f} ms/step")
    st_ms = capi.baryons_udsc_profile(*v32, UDSC_SRC, T, moms, capi.MODE_MIXED, capi.MOMENTUM_SPACE, capi.F64,
                                      out_mx, stream=stream)
    cmac, nops = udsc_work_per_site()
    res = {"metric": "udsc baryon 2pt sites/s (20 flavour products, 200 operators), mixed precision",
           "value": nsites / (ms_mx / 1e3), "unit": "sites/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_mx,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": proj_dtype(st_ms, 6),
           "data": "synthetic",
           "config": {"workload": "udsc: lattice 32^3x64, one source, u/d/s/c propagators, all 20 flavour "
                                  "products (200 operators), 64 time slices, |n|^2<=9 (123 momenta)",
                      "lattice": L, "sources": 1, "sites": nsites, "operators": nops, "nmom": len(moms),
                      "parallelism": "single GPU"},
           "stages_ms": st_ms}
    if not args.no_fp64:
        plain()
        ms_64, _ = timed(torch, stream, plain, max(1, args.steps // 2))
        st64 = capi.baryons_udsc_profile(*v64, UDSC_SRC, T, moms, capi.MODE_F64, capi.MOMENTUM_SPACE, capi.F64,
                                         out64, stream=stream)
        torch.cuda.synchronize()
        a = out_mx.cpu().numpy()
        b = out64.cpu().numpy()
        A = (a[:, 0] + 1j * a[:, 1]).reshape(nops, T, len(moms))
        B = (b[:, 0] + 1j * b[:, 1]).reshape(nops, T, len(moms))
        den = np.abs(B).max(axis=2, keepdims=True)
        E = np.abs(A - B) / np.where(den == 0, 1.0, den)
        res["fp64_plain"] = {"value": nsites / (ms_64 / 1e3), "unit": "sites/s", "ms_per_step": ms_64,
                             "stages_ms": st64}
        res["speedup_vs_fp64"] = res["value"] / res["fp64_plain"]["value"]
        res["rel_err_vs_fp64"] = {"max": float(E.max()), "median": float(np.median(E)),
                                  "definition": "max over operators, time slices and momenta of "
                                                "|C_mixed - C_fp64| / max_momenta |C_fp64|"}
    dom = max(st_ms, key=st_ms.get)
    prop_bytes = 20 * 3 * 144 * 8
    if dom.startswith("contract"):
        r = roof(dom, st_ms[dom] / 1e3, nsites, 8 * cmac, prop_bytes + nops * 8, "valu", FP32_VALU_PEAK_TFS,
                 "udsc", f"{8 * cmac} implemented FP32 flop over the 20 products; 20 products x 3 x 1152 B "
                         f"FP32 propagator reads + {nops} x 8 B site correlator writes per site")
    else:
        r = roofline(st_ms, nsites, len(moms), nops, 6, 0, 0, "udsc")
    res["roofline"] = r
    if not args.no_cpu:
        log("cpu baseline")
        res["cpu_baseline"] = cpu_baseline_udsc()
    return res


TETRA_PATTERNS = [((0, 2, 1, 3, 2, 0, 3, 1), 0), ((0, 2, 1, 3, 3, 0, 2, 1), 1), ((0, 3, 1, 2, 2, 0, 3, 1), 1),
                  ((0, 3, 1, 2, 3, 0, 2, 1), 0)]
TETRA_WEIGHTS = [9600, 9600, 3072, 3072]   # pattern counts of the reference's tables


def tetra_table(nops=16, nterm=768):
    """synthetic table with the reference's operator count, term count and
    colour-pattern proportions"""
    rng = np.random.default_rng(SEED)
    w = np.array(TETRA_WEIGHTS, dtype=float) / sum(TETRA_WEIGHTS)
    n = nops * nterm
    pat = rng.choice(4, size=n, p=w)
    cols = np.array([TETRA_PATTERNS[k][0] for k in pat], dtype=np.int16)
    idxs = rng.integers(0, 4, size=(n, 8)).astype(np.int16)
    vals = np.stack([rng.choice([-1.0, 1.0, 0.5], size=n), rng.choice([0.0, 1.0, -1.0], size=n)], -1).astype(np.float32)
    counts = np.full(nops, nterm, dtype=np.int16)
    direct = int(sum(1 for k in pat if TETRA_PATTERNS[k][1] == 0))
    return idxs, cols, vals, counts, direct


def cpu_baseline_tetra(tab, budget_s=12.0):
    """Oracle (the reference's 81-term colour sums per term, FP64, + per-site
    DFT, OpenMP) on an 8^3 spatial lattice: the per-site work does not depend
    on the lattice size"""
    from oracle import oracle as O
    nth = baseline_threads()
    Ls = [8, 8, 8, 16]
    V3 = Ls[0] * Ls[1] * Ls[2]
    P = [O.synth_prop(Ls, SEED, i, 0, lam, nthreads=nth) for i, lam in enumerate((0.5, 0.5, 1.2, 1.2))]
    P[1] = P[0]
    P[3] = P[2]
    idxs, cols, vals, counts, _ = tab
    v = vals[:, 0].astype(np.float64) + 1j * vals[:, 1]
    t0 = time.perf_counter()
    O.tetra_correlator(Ls, [1, 2, 3, 0], 1, Q2MAX4, P, idxs, cols, v, counts, nthreads=nth)
    used = time.perf_counter() - t0
    return {"value": V3 / used, "unit": "sites/s", "cores": nth, "kind": "port", **cpu_info(),
            "sample": f"1 time slice x {V3} sites of an 8^3x16 lattice, 16 operators x 768 terms, 123 momenta, "
                      f"FP64 81-term colour sums + per-site DFT, {used:.1f} s"}


def run_tetra(args, torch, capi, dev, stream):
    capi.init(L, 0)
    V3 = L[0] * L[1] * L[2]
    T = L[3]
    src = [3, 5, 7, 10]
    moms = capi.momlist(Q2MAX4)
    log("synthesising light and heavy FP64 propagators and FP32 copies")
    keep, v64, v32 = [], [], []
    for fl, lam in enumerate((LAMBDA, LAMBDA_HEAVY)):
        t64 = torch.empty((144, T * V3, 2), dtype=torch.float64, device=dev)
        v = capi.plg_prop(t64.data_ptr(), capi.F64, 0, T, 0, T * V3)
        capi.synth_prop(v, SEED, 60 + fl, src[3], lam)
        t32 = torch.empty((144, T * V3, 2), dtype=torch.float32, device=dev)
        e32 = torch.empty(T, dtype=torch.int32, device=dev)
        w = capi.plg_prop(t32.data_ptr(), capi.F32, 0, T, 0, T * V3, e32.data_ptr())
        capi.cast_prop(w, v)
        keep += [t64, t32, e32]
        v64.append(v)
        v32.append(w)
    torch.cuda.synchronize()
    tab = tetra_table()
    idxs, cols, vals, counts, ndirect = tab
    nops = len(counts)
    out_mx = torch.zeros((nops * T * len(moms), 2), dtype=torch.float64, device=dev)
    out64 = torch.zeros_like(out_mx)
    views32 = [v32[0], v32[0], v32[1], v32[1]]
    views64 = [v64[0], v64[0], v64[1], v64[1]]

    def mixed():
        capi.tetraquark_table(*views32, idxs, cols, vals, counts, src, T, moms, capi.MODE_MIXED,
                              capi.MOMENTUM_SPACE, capi.F64, out_mx, stream=stream)

    def plain():
        capi.tetraquark_table(*views64, idxs, cols, vals, counts, src, T, moms, capi.MODE_F64,
                              capi.MOMENTUM_SPACE, capi.F64, out64, stream=stream)

    nsites = V3 * T
    log(f"mixed: {nsites} sites, {len(idxs)} terms; warmup {args.warmup}")
    for _ in range(args.warmup):
        mixed()
    ms_mx, _ = timed(torch, stream, mixed, args.steps)
    log(f"mixed: {ms_mx:.3f} ms/step")
    # implemented complex MACs per site: the spin-pair tables (256 entries of
    # <B1,B3>, B3^dag B1, <B2,B4>, B4^dag B2: 2 x 256 x 36), then 1 per direct
    # term, 9 per crossed term and 1 per coefficient
    cmac = 2 * 256 * 36 + ndirect + 9 * (len(idxs) - ndirect) + len(idxs)
    # the colour sums of the table terms as the reference writes them: 81
    # products of four propagator entries (3 complex multiplications) per term
    cmac_ref = 81 * 3 * len(idxs)
    res = {"metric": "tetraquark table 2pt sites/s (16 operators x 768 terms), mixed precision",
           "value": nsites / (ms_mx / 1e3), "unit": "sites/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms_mx, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32 contraction + autotuned projection (fp64 separable sums "
                                          "or bf16x6 MFMA GEMM), fp64 sums",
           "data": "synthetic",
           "config": {"workload": "tetra: lattice 32^3x64, one source, light/light/heavy/heavy propagators, "
                                  "synthetic 16 x 768-term table in the reference's colour patterns, 64 time "
                                  "slices, |n|^2<=9 (123 momenta)",
                      "lattice": L, "sites": nsites, "operators": nops, "terms": int(len(idxs)),
                      "nmom": len(moms), "parallelism": "single GPU"},
           "contraction_valu": {"flops_per_site": 8 * cmac, "reference_body_flops_per_site": 8 * cmac_ref,
                                "achieved_upper_bound_TFLOPs": nsites * 8 * cmac / (ms_mx / 1e3) / 1e12,
                                "note": "whole step (contraction + projection) in the denominator; the kernel "
                                        "reads the spin-pair tables from LDS (about 1 MB per site), which bounds it"}}
    if not args.no_fp64:
        plain()
        ms_64, _ = timed(torch, stream, plain, max(1, args.steps // 2))
        torch.cuda.synchronize()
        a = out_mx.cpu().numpy()
        b = out64.cpu().numpy()
        A = (a[:, 0] + 1j * a[:, 1]).reshape(nops, T, len(moms))
        B = (b[:, 0] + 1j * b[:, 1]).reshape(nops, T, len(moms))
        den = np.abs(B).max(axis=2, keepdims=True)
        E = np.abs(A - B) / np.where(den == 0, 1.0, den)
        res["fp64_plain"] = {"value": nsites / (ms_64 / 1e3), "unit": "sites/s", "ms_per_step": ms_64}
        res["speedup_vs_fp64"] = res["value"] / res["fp64_plain"]["value"]
        res["rel_err_vs_fp64"] = {"max": float(E.max()), "median": float(np.median(E))}
    if not args.no_cpu:
        log("cpu baseline")
        res["cpu_baseline"] = cpu_baseline_tetra(tab)
    return res


def learned_flow(torch, capi, dev, stream, Lat=L, nsrc=256, split=(32, 32), steps=2000, seed=SEED + 5,
                 log_every=64):
    """The learned workload on lattice Lat with nsrc sources, propagators a
    shared seeded base plus a 5% seeded per-source perturbation:
      exact:  the plain FP64 path (16 mesons + nucleon, |n|^2 <= 16) of every
              source (the labels, the truth the estimators are judged by, and
              the fallback);
      cheap:  the mixed path at n = 0 only (16 mesons + nucleon in FP32
              contraction), every source;
      learned flow: cheap inputs of all sources + exact labels of the first
              split[0] + split[1] sources, the network trained on the first
              split[0] (its time slices are the samples), bias-corrected on
              the next split[1], predicting the rest.
    Features and targets are in source-relative time; the targets have the
    source phase removed.  Returns (result dict, arrays)."""
    from plegma_amd import learned
    capi.init(Lat, 0)
    V3, T = Lat[0] * Lat[1] * Lat[2], Lat[3]
    rng = np.random.default_rng(seed)
    srcs = [[int(rng.integers(0, Lat[i])) for i in range(4)] for _ in range(nsrc)]
    moms = capi.momlist(Q2MAX)
    kind, mask = capi.TWOPT, capi.G_NUCL
    t64 = torch.empty((144, T * V3, 2), dtype=torch.float64, device=dev)
    v = capi.plg_prop(t64.data_ptr(), capi.F64, 0, T, 0, T * V3)
    t32 = torch.empty((144, T * V3, 2), dtype=torch.float32, device=dev)
    e32 = torch.empty(T, dtype=torch.int32, device=dev)
    w = capi.plg_prop(t32.data_ptr(), capi.F32, 0, T, 0, T * V3, e32.data_ptr())
    n = capi.correlator_size(kind, mask, srcs[0], T, len(moms))
    out = torch.empty((n, 2), dtype=torch.float64, device=dev)
    n0 = capi.correlator_size(kind, mask, srcs[0], T, 1)
    out0 = torch.empty((n0, 2), dtype=torch.float64, device=dev)

    def ev_ms(fn):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1)

    # one untimed call of each path first: the plans' tuning runs once per
    # geometry (and is cached on disk), as in the other workloads' warm-up
    capi.synth_prop_perturbed(v, seed, 1000, 1, 0.05, srcs[0][3], LAMBDA, stream)
    capi.correlator(kind, mask, v, v, srcs[0], T, moms, capi.MODE_F64, capi.MOMENTUM_SPACE, capi.F64, out, stream)
    capi.cast_prop(w, v, stream)
    capi.correlator(kind, mask, w, w, srcs[0], T, [(0, 0, 0)], capi.MODE_MIXED, capi.MOMENTUM_SPACE, capi.F64,
                    out0, stream)
    torch.cuda.synchronize()
    X, Y, t_exact, t_cheap = [], [], 0.0, 0.0
    for s, src in enumerate(srcs):
        capi.synth_prop_perturbed(v, seed, 1000, 1 + s, 0.05, src[3], LAMBDA, stream)
        t_exact += ev_ms(lambda: capi.correlator(kind, mask, v, v, src, T, moms, capi.MODE_F64,
                                                 capi.MOMENTUM_SPACE, capi.F64, out, stream))
        capi.cast_prop(w, v, stream)
        t_cheap += ev_ms(lambda: capi.correlator(kind, mask, w, w, src, T, [(0, 0, 0)], capi.MODE_MIXED,
                                                 capi.MOMENTUM_SPACE, capi.F64, out0, stream))
        times = [capi.row_time(src, T, j) for j in range(T)]
        h = out.cpu().numpy()
        h0 = out0.cpu().numpy()
        x, a = learned.cheap_features(h0[:, 0] + 1j * h0[:, 1], times, src, T)
        X.append(learned.RowwiseNet.pack(x, a))
        Y.append(learned.shell_targets(h[:, 0] + 1j * h[:, 1], moms, times, src, Lat).reshape(-1))
        if (s + 1) % log_every == 0:
            log(f"  {s + 1} sources")
    del t64, t32, e32
    X, Y = np.array(X), np.array(Y)
    ntr, nbc = split
    nlab = ntr + nbc
    npred = nsrc - nlab
    tr, bc, pr = slice(0, ntr), slice(ntr, nlab), slice(nlab, nsrc)
    log(f"training the network on the {ntr * T} time slices of {ntr} sources")
    net = learned.RowwiseNet(T, hidden=(256, 256), seed=SEED, max_batch=max(4096, ntr * T))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    final_loss = net.fit(X[tr], Y[tr], steps=steps, lr=1e-3)
    est = learned.bias_corrected(net, X[bc], Y[bc], X[pr])
    torch.cuda.synchronize()
    t_ml = (time.perf_counter() - t0) * 1e3
    net.close()
    gain_est = learned.expected_gain(est, Y[:nlab], nlab, nbc, npred)
    # the error of each estimator of the mean over all sources, per output
    # (time slice, shell) relative to the spread of the output over sources
    scale = Y.std(axis=0)
    scale = np.where(scale == 0, 1.0, scale)
    truth = Y.mean(axis=0)
    m_ml = (Y[:nlab].sum(axis=0) + est.pred.sum(axis=0)) / nsrc
    m_lab = Y[:nlab].mean(axis=0)
    e_ml = float(np.mean(((m_ml - truth) / scale) ** 2))
    e_lab = float(np.mean(((m_lab - truth) / scale) ** 2))
    per = np.abs(est.pred - Y[pr]) / scale                        # [source][time slice x shell]
    # the plain FP64 path is the fallback: when the held-out sources say the
    # bias-corrected estimator would not beat the labelled sources alone, the
    # predicted sources are computed exactly
    fb = learned.with_fallback(est, lambda: Y[pr], gain=gain_est, min_gain=1.0)
    cost_exact = t_exact
    cost_ml = t_cheap + t_exact * nlab / nsrc + t_ml + (t_exact * npred / nsrc if fb.used_fallback else 0.0)
    nsites = nsrc * V3 * T
    res = {"exact_fp64_ms_all_sources": cost_exact,
           "ml_flow_ms": {f"cheap_inputs_{nsrc}": t_cheap, f"exact_labels_{nlab}": t_exact * nlab / nsrc,
                          "train_and_predict": t_ml},
           "sites": nsites, "flow_ms": cost_ml,
           "network": f"17-256-256-11 per time slice (weights shared over the {T} time slices of a source), "
                      f"BF16 weights, FP32 accumulation, {steps} full-batch Adam steps on {ntr * T} samples",
           "train_loss_standardised": final_loss,
           "pred_err_per_source": {"median": float(np.median(per)), "p90": float(np.quantile(per, 0.9)),
                                   "max": float(per.max()),
                                   "definition": "|pred - exact| / spread over sources, per source, time "
                                                 "slice and shell"},
           "heldout_rel_resid_median": float(np.median(est.rel_error_estimate())),
           "expected_gain_from_heldout": gain_est, "fallback_used": bool(fb.used_fallback),
           "fallback_rule": "compute the predicted sources exactly unless the held-out sources put the "
                            "bias-corrected estimator's expected gain above 1",
           "speedup_vs_exact_all": cost_exact / cost_ml,
           "ensemble_mean_sq_err": {"bias_corrected": e_ml, "labelled_only": e_lab,
                                    "gain": e_lab / e_ml if e_ml > 0 else None,
                                    "definition": "mean over (time slice, shell) of ((estimate - mean over "
                                                  "all sources) / spread over sources)^2"}}
    return res, {"X": X, "Y": Y, "est": est}


def run_ft(args, torch, capi, dev, stream):
    """the standalone projection of a field (PLEGMA_FT::apply,
    lib/PLEGMA_FT.cu): a 4D field of 16 complex components on 32^3 x 64,
    all momenta with |n|^2 <= 16, FP32 and FP64 fields, the separable sums
    (the default) and the phase-matrix GEMM (PLG_PROJ_SEP=0)"""
    capi.init(L, 0)
    V3, T, dof = L[0] * L[1] * L[2], L[3], 16
    moms = capi.momlist(Q2MAX)
    moms_i32 = np.ascontiguousarray(moms, dtype=np.int32)     # built once, passed as is
    g = torch.Generator(device=dev)
    g.manual_seed(SEED)
    f64 = torch.randn((dof, T, V3, 2), dtype=torch.float64, device=dev, generator=g)
    f32 = f64.to(torch.float32)
    out = {k: torch.zeros((T, dof, len(moms), 2), dtype=torch.float64, device=dev) for k in ("32", "64")}
    ms = {}
    for sep in ("1", "0"):
        os.environ["PLG_PROJ_SEP"] = sep
        try:
            for prec, fld in (("32", f32), ("64", f64)):
                o = out[prec] if sep == "1" else torch.zeros_like(out[prec])
                fn = lambda: capi.ft_apply(fld, capi.F32 if prec == "32" else capi.F64, dof, T, moms_i32, -1, o,
                                           stream=stream)
                for _ in range(args.warmup):
                    fn()
                ms[(sep, prec)], _ = timed(torch, stream, fn, args.steps)
                log(f"ft {'separable' if sep == '1' else 'GEMM'} FP{prec}: {ms[(sep, prec)]:.3f} ms")
                if sep == "0":
                    ref = out[prec]
                    ms[("diff", prec)] = ((o - ref).abs().max() / ref.abs().max()).item()
        finally:
            del os.environ["PLG_PROJ_SEP"]
    a, b = out["32"].cpu().numpy(), out["64"].cpu().numpy()
    A, B = a[..., 0] + 1j * a[..., 1], b[..., 0] + 1j * b[..., 1]
    nsites = V3 * T
    t32 = ms[("1", "32")] / 1e3
    res = {"metric": "standalone FT sites/s (16 components, 257 momenta), FP32 field", "value": nsites / t32,
           "unit": "sites/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": ms[("1", "32")], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32 field, fp64 separable sums", "data": "synthetic",
           "config": {"workload": "ft: lattice 32^3x64, a 4D field of 16 complex components (torch.randn), "
                                  "|n|^2<=16 (257 momenta), sign -1",
                      "lattice": L, "components": dof, "nmom": len(moms), "parallelism": "single GPU"},
           "fp64_field": {"ms_per_step": ms[("1", "64")], "value": nsites / (ms[("1", "64")] / 1e3)},
           "gemm": {"fp32_ms": ms[("0", "32")], "fp64_ms": ms[("0", "64")],
                    "max_diff_vs_separable_rel": {"fp32": ms[("diff", "32")], "fp64": ms[("diff", "64")]}},
           "rel_err_fp32_vs_fp64": float(np.abs(A - B).max() / np.abs(B).max()),
           "roofline": roof("ft_apply_f32", t32, nsites, 0, dof * 8, "hbm", 1.0, "ft",
                            f"{dof * 8} B FP32 field read per site; whole call (the sums over x and y, "
                            "over z, the output layout, the stream synchronisation)")}
    if not args.no_cpu:
        # the oracle's per-site DFT of one time slice, all components
        import oracle.oracle as O
        nth = len(os.sched_getaffinity(0))
        t0, n = time.perf_counter(), 0
        while n < T and (n == 0 or time.perf_counter() - t0 < 10.0):
            h = f64[:, n].cpu().numpy()
            O.dft(L, [0, 0, 0], Q2MAX, np.ascontiguousarray((h[..., 0] + 1j * h[..., 1]).T), nth)
            n += 1
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": n * V3 / dt, "unit": "sites/s", "cores": nth, "kind": "port",
                               "sample": f"{n} time slices of {V3} sites x {dof} components, {len(moms)} momenta, "
                                         f"FP64 per-site DFT (oracle), {dt:.1f} s"}
    return res


def run_learned(args, torch, capi, dev, stream):
    """32^3 x 64, 256 sources (config of the learned part); see learned_flow"""
    log("learned flow: exact FP64 and cheap mixed two-point results of 256 sources")
    r, _ = learned_flow(torch, capi, dev, stream)
    res = {"metric": METRIC, "value": r["sites"] / (r["flow_ms"] / 1e3), "unit": "sites/s", "n_gpus": 1,
           "steps": 1, "warmup": 0, "ms_per_step": r["flow_ms"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None,
           "dtype": "f64 two-point path (labels); f32 contraction + autotuned projection at n = 0 (inputs); "
                    "bf16 MLP (fp32 accumulation)",
           "data": "synthetic",
           "config": {"workload": "learned: lattice 32^3x64, 256 sources, propagators base + 5% "
                                  "perturbation, MLP 17-256-256-11 per time slice, train 32 / bias-correct 32 / "
                                  "predict 192, nucleon at |n|^2 in 4..16 (11 shells)",
                      "lattice": L, "sources": 256, "sites": r["sites"], "parallelism": "single GPU"},
           "learned": r}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--fp64-steps", type=int, default=None)
    ap.add_argument("--split", type=int, default=6)
    ap.add_argument("--workload", choices=["batching", "largest", "learned", "udsc", "tetra", "ft"],
                    default="batching")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--calib-copy", action="store_true",
                    help="first run a plain 1 GiB device copy (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              f"(sources and time slices would shard over GPUs without communication)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from plegma_amd import capi

    assert torch.cuda.is_available(), "bench needs a HIP device"
    dev = torch.device("cuda:0")
    if args.calib_copy:
        # 1 GiB read + 1 GiB written by one copy kernel: the reference point
        # for the FETCH_SIZE / WRITE_SIZE counters of the other kernels
        a = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    stream = torch.cuda.Stream()
    run = {"largest": run_largest, "learned": run_learned, "udsc": run_udsc, "tetra": run_tetra,
           "ft": run_ft}.get(args.workload, run_batching)
    res = run(args, torch, capi, dev, stream)
    capi.finalize()
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
