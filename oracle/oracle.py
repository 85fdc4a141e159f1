"""ctypes front end of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product path never imports this module.

Restates PLEGMA's 2pt path in FP64 (see plegma_oracle.h for the reference
files each function follows).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

MESONS, MESONS16, MESONS_NEW, MESONS_ALL, BARYONS, MESON_MESON, TWOPT = range(7)
G_NUCL = 1 << 0
G_DELTA12 = (1 << 4, 1 << 5, 1 << 6)
G_DELTA32 = (1 << 7, 1 << 8, 1 << 9)


class OrcProp(C.Structure):
    _fields_ = [("data", C.c_void_p), ("comp_stride", C.c_int64),
                ("t_first", C.c_int), ("nt", C.c_int)]


def build(force: bool = False) -> str:
    """Compile liboracle.so with its Makefile (gcc, OpenMP)."""
    if force or not os.path.exists(_LIB):
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=_HERE, check=True)
    return _LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        i32p = C.POINTER(C.c_int)
        f64p = C.c_void_p
        L.orc_nmom.argtypes = [C.c_int]
        L.orc_momlist.argtypes = [C.c_int, i32p]
        L.orc_window.argtypes = [i32p, i32p, C.c_int, i32p, i32p, i32p]
        L.orc_row_time.argtypes = [i32p, i32p, C.c_int, C.c_int]
        L.orc_kind_nch.argtypes = [C.c_int, C.c_int, i32p, i32p]
        for n in ("orc_site_mesons10", "orc_site_mesons16", "orc_site_mesons_new",
                  "orc_site_mesons_all", "orc_site_nton", "orc_site_meson_meson"):
            getattr(L, n).argtypes = [f64p, f64p, f64p]
            getattr(L, n).restype = None
        for n in ("orc_site_delta32", "orc_site_delta12"):
            getattr(L, n).argtypes = [C.c_int, f64p, f64p, f64p]
            getattr(L, n).restype = None
        L.orc_correlator.argtypes = [C.c_int, C.c_int, i32p, i32p, C.c_int, C.c_int,
                                     C.POINTER(OrcProp), C.POINTER(OrcProp), f64p,
                                     i32p, C.c_int, C.c_int]
        L.orc_site_correlator.argtypes = [C.c_int, C.c_int, i32p, C.c_int,
                                          C.POINTER(OrcProp), C.POINTER(OrcProp), f64p,
                                          C.c_int]
        L.orc_dft.argtypes = [i32p, i32p, C.c_int, f64p, C.c_int, f64p, C.c_int]
        L.orc_dft.restype = None
        L.orc_synth_prop.argtypes = [i32p, C.c_uint64, C.c_uint32, C.c_int, C.c_double,
                                     f64p, C.c_int64, C.c_int, C.c_int, C.c_int]
        L.orc_synth_prop.restype = None
        L.orc_normal_pair.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32,
                                      C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_normal_pair.restype = None
        L.orc_philox.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                 C.POINTER(C.c_uint32)]
        L.orc_philox.restype = None
        L.orc_meson_table.argtypes = [C.c_int, i32p, f64p]
        L.orc_diquark_table.argtypes = [C.c_int, i32p, f64p]
        L.orc_gamma.argtypes = [C.c_int, f64p]
        L.orc_rotate_site.argtypes = [f64p, C.c_int, f64p]
        L.orc_prop_rotate.argtypes = [f64p, C.c_int64, C.c_int, C.c_int64, C.c_int]
        L.orc_prop_boundaries.argtypes = [f64p, C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int64, C.c_int]
        L.orc_gamma.restype = None
        L.orc_site_udsc.argtypes = [f64p, f64p, f64p, i32p, f64p, i32p, i32p, C.c_int, f64p]
        L.orc_site_udsc.restype = None
        L.orc_udsc_field.argtypes = [i32p, C.c_int, C.POINTER(OrcProp), C.POINTER(OrcProp), C.POINTER(OrcProp),
                                     i32p, f64p, i32p, i32p, C.c_int, f64p, C.c_int]
        L.orc_site_tetra.argtypes = [f64p, f64p, f64p, f64p, i32p, i32p, f64p, i32p, i32p, C.c_int, f64p]
        L.orc_site_tetra.restype = None
        L.orc_tetra_field.argtypes = [i32p, C.c_int] + [C.POINTER(OrcProp)] * 4 + [i32p, i32p, f64p, i32p, i32p,
                                                                                   C.c_int, f64p, C.c_int]
        _lib = L
    return _lib


def _ivec(v):
    a = (C.c_int * len(v))(*[int(x) for x in v])
    return a


def _ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def momlist(q2max: int) -> np.ndarray:
    n = lib().orc_nmom(q2max)
    out = (C.c_int * (3 * n))()
    lib().orc_momlist(q2max, out)
    return np.array(out, dtype=np.int32).reshape(n, 3)


def window(L, src, totalT):
    s, e, lt = C.c_int(), C.c_int(), C.c_int()
    lib().orc_window(_ivec(L), _ivec(src), totalT, C.byref(s), C.byref(e), C.byref(lt))
    return s.value, e.value, lt.value


def row_time(L, src, totalT, j):
    return lib().orc_row_time(_ivec(L), _ivec(src), totalT, j)


def kind_shape(kind, mask=G_NUCL):
    nd, ng = C.c_int(), C.c_int()
    nch = lib().orc_kind_nch(kind, mask, C.byref(nd), C.byref(ng))
    return ng.value, nd.value, nch


def gamma(op: int) -> np.ndarray:
    m = np.zeros(32)
    lib().orc_gamma(op, _ptr(m))
    return (m[0::2] + 1j * m[1::2]).reshape(4, 4)


def meson_table(op: int):
    idx = (C.c_int * 64)()
    val = np.zeros(32)
    lib().orc_meson_table(op, idx, _ptr(val))
    return np.array(idx).reshape(16, 4), (val[0::2] + 1j * val[1::2])


def diquark_table(which: int):
    idx = (C.c_int * 64)()
    val = np.zeros(32)
    lib().orc_diquark_table(which, idx, _ptr(val))
    return np.array(idx).reshape(16, 4), (val[0::2] + 1j * val[1::2])


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox(c, k, o)
    return list(o)


def normal_pair(seed, stream, site, comp):
    re, im = C.c_double(), C.c_double()
    lib().orc_normal_pair(seed, stream, site, comp, C.byref(re), C.byref(im))
    return re.value, im.value


def site_body(name: str, S1: np.ndarray, S2: np.ndarray, k: int | None = None) -> np.ndarray:
    """Run one per-site body on complex [4,4,3,3] arrays; returns complex accum[]."""
    sizes = {"mesons10": 20, "mesons16": 32, "mesons_new": 10, "mesons_all": 34,
             "nton": 32, "delta32": 32, "delta12": 32, "meson_meson": 20}
    a = np.ascontiguousarray(np.stack([S1.real, S1.imag], -1).reshape(-1))
    b = np.ascontiguousarray(np.stack([S2.real, S2.imag], -1).reshape(-1))
    out = np.zeros(2 * sizes[name])
    f = getattr(lib(), "orc_site_" + name)
    if k is None:
        f(_ptr(a), _ptr(b), _ptr(out))
    else:
        f(k, _ptr(a), _ptr(b), _ptr(out))
    return out[0::2] + 1j * out[1::2]


def _prop(data: np.ndarray, comp_stride: int, t_first: int = 0, nt: int | None = None,
          T: int | None = None):
    return OrcProp(data.ctypes.data, comp_stride, t_first, nt if nt is not None else T)


def correlator(kind, L, src, totalT, q2max, prop1, prop2, mask=G_NUCL, rows=None,
               nthreads=0, view1=None, view2=None):
    """H_elem of the correlator (complex array [g][d][j][m][c]).

    prop1/prop2: float64 arrays in PLEGMA SoA layout (144, planes*V3, 2) flattened.
    view*: optional (comp_stride, t_first, nt)."""
    ng, nd, nch = kind_shape(kind, mask)
    s, e, lt = window(L, src, totalT)
    nmom = lib().orc_nmom(q2max)
    V3 = L[0] * L[1] * L[2]
    H = np.zeros(ng * nd * lt * nmom * nch * 2)
    v1 = view1 or (L[3] * V3, 0, L[3])
    v2 = view2 or (L[3] * V3, 0, L[3])
    p1 = OrcProp(prop1.ctypes.data, v1[0], v1[1], v1[2])
    p2 = OrcProp(prop2.ctypes.data, v2[0], v2[1], v2[2])
    rr = None if rows is None else _ivec(rows)
    rc = lib().orc_correlator(kind, mask, _ivec(L), _ivec(src), totalT, q2max,
                              C.byref(p1), C.byref(p2), _ptr(H), rr,
                              0 if rows is None else len(rows), nthreads)
    if rc != 0:
        raise RuntimeError(f"orc_correlator failed: {rc}")
    H = H.reshape(ng, nd, lt, nmom, nch, 2)
    return H[..., 0] + 1j * H[..., 1]


def site_correlator(kind, L, t, prop1, prop2, mask=G_NUCL, nthreads=0,
                    view1=None, view2=None):
    ng, nd, nch = kind_shape(kind, mask)
    V3 = L[0] * L[1] * L[2]
    out = np.zeros(V3 * ng * nd * nch * 2)
    v1 = view1 or (L[3] * V3, 0, L[3])
    v2 = view2 or (L[3] * V3, 0, L[3])
    p1 = OrcProp(prop1.ctypes.data, v1[0], v1[1], v1[2])
    p2 = OrcProp(prop2.ctypes.data, v2[0], v2[1], v2[2])
    rc = lib().orc_site_correlator(kind, mask, _ivec(L), t, C.byref(p1), C.byref(p2),
                                   _ptr(out), nthreads)
    if rc != 0:
        raise RuntimeError(f"orc_site_correlator failed: {rc}")
    out = out.reshape(V3, nd * ng * nch, 2)
    return out[..., 0] + 1j * out[..., 1]


def dft(L, src3, q2max, f: np.ndarray, nthreads=0) -> np.ndarray:
    """f: complex [V3, ncomp] -> [nmom, ncomp]"""
    V3, ncomp = f.shape
    fr = np.ascontiguousarray(np.stack([f.real, f.imag], -1).reshape(-1))
    nmom = lib().orc_nmom(q2max)
    out = np.zeros(nmom * ncomp * 2)
    lib().orc_dft(_ivec(L), _ivec(src3), q2max, _ptr(fr), ncomp, _ptr(out), nthreads)
    out = out.reshape(nmom, ncomp, 2)
    return out[..., 0] + 1j * out[..., 1]


def synth_prop(L, seed, stream, t0, lam, t_first=0, nt=None, comp_stride=None,
               nthreads=0) -> np.ndarray:
    """Seeded synthetic propagator, float64 SoA array of 144*comp_stride*2."""
    V3 = L[0] * L[1] * L[2]
    nt = L[3] if nt is None else nt
    cs = nt * V3 if comp_stride is None else comp_stride
    data = np.zeros(144 * cs * 2)
    lib().orc_synth_prop(_ivec(L), seed, stream, t0, lam, _ptr(data), cs, t_first, nt,
                         nthreads)
    return data


def prop_to_sites(prop: np.ndarray, L, t, comp_stride=None, t_first=0) -> np.ndarray:
    """complex [V3, 4,4,3,3] view of one time slice of a SoA propagator"""
    V3 = L[0] * L[1] * L[2]
    cs = L[3] * V3 if comp_stride is None else comp_stride
    plane = (t - t_first) % L[3]
    p = prop.reshape(144, cs, 2)[:, plane * V3:(plane + 1) * V3, :]
    c = p[..., 0] + 1j * p[..., 1]
    return np.ascontiguousarray(c.T).reshape(V3, 4, 4, 3, 3)


def rotate_site(S, sign):
    """rotation to the physical basis of one site: S complex [144] -> [144]"""
    a = np.ascontiguousarray(np.stack([np.real(S), np.imag(S)], -1), dtype=np.float64)
    out = np.empty_like(a)
    lib().orc_rotate_site(a.ctypes.data_as(C.POINTER(C.c_double)), int(sign),
                          out.ctypes.data_as(C.POINTER(C.c_double)))
    return out[..., 0] + 1j * out[..., 1]


def prop_rotate(P, L, sign):
    """in place on a float64 propagator array [144, nt*V3, 2]"""
    assert P.dtype == np.float64 and P.flags.c_contiguous
    V3 = L[0] * L[1] * L[2]
    lib().orc_prop_rotate(P.ctypes.data_as(C.POINTER(C.c_double)), P.shape[1], P.shape[1] // V3, V3, int(sign))


def prop_boundaries(P, L, t0, t_first=0):
    """in place on a float64 propagator array [144, nt*V3, 2]"""
    assert P.dtype == np.float64 and P.flags.c_contiguous
    V3 = L[0] * L[1] * L[2]
    lib().orc_prop_boundaries(P.ctypes.data_as(C.POINTER(C.c_double)), P.shape[1], t_first,
                              P.shape[1] // V3, L[3], V3, int(t0))


def udsc_tables(prod_tables):
    """flat arrays of one flavour product's operator tables [(name, {m6: v})]"""
    idx, val, first, count = [], [], [], []
    for _, t in prod_tables:
        first.append(len(idx) // 6)
        count.append(len(t))
        for m in sorted(t):
            idx.extend(m)
            val.extend([t[m].real, t[m].imag])
    return (np.array(idx, dtype=np.int32), np.array(val, dtype=np.float64), np.array(first, dtype=np.int32),
            np.array(count, dtype=np.int32))


def udsc_site(S1, S2, S3, prod_tables):
    """operators of one product on one site; S*: complex [4,4,3,3]"""
    idx, val, first, count = udsc_tables(prod_tables)
    raw = [np.ascontiguousarray(np.stack([S.real, S.imag], -1).reshape(-1)) for S in (S1, S2, S3)]
    out = np.zeros(2 * len(count))
    lib().orc_site_udsc(_ptr(raw[0]), _ptr(raw[1]), _ptr(raw[2]), idx.ctypes.data_as(C.POINTER(C.c_int)),
                        _ptr(val), first.ctypes.data_as(C.POINTER(C.c_int)),
                        count.ctypes.data_as(C.POINTER(C.c_int)), len(count), _ptr(out))
    return out[0::2] + 1j * out[1::2]


def udsc_correlator(L, src, totalT, q2max, props, products, only_st=False, only_ch=False, space="momentum",
                    nthreads=0):
    """contractBaryonsUDSC (lib/PLEGMA_Correlator.cu:136-188) on the CPU.

    props: flavour letter ('u', 'd', 's', 'c') -> float64 SoA propagator
    (144, T*V3, 2) flattened, absent flavours missing; products: the
    [(letters, [(name, {m6: v})])] term tables.  Returns (names, H) with
    H[d][j][m] (momentum space) or H[d][j][x] (position space)."""
    V3 = L[0] * L[1] * L[2]
    s, e, lt = window(L, src, totalT)
    names, blocks = [], []
    for letters, tabs in products:
        if any(f not in props for f in letters):
            continue
        if only_st and "s" not in letters:
            continue
        if only_ch and "c" not in letters:
            continue
        idx, val, first, count = udsc_tables(tabs)
        P = [OrcProp(props[f].ctypes.data, L[3] * V3, 0, L[3]) for f in letters]
        nops = len(count)
        H = np.zeros((nops, lt) + ((len(momlist(q2max)),) if space == "momentum" else (V3,)), dtype=complex)
        for j in range(lt):
            t = row_time(L, src, totalT, j)
            f = np.zeros(V3 * nops * 2)
            rc = lib().orc_udsc_field(_ivec(L), t, C.byref(P[0]), C.byref(P[1]), C.byref(P[2]),
                                      idx.ctypes.data_as(C.POINTER(C.c_int)), _ptr(val),
                                      first.ctypes.data_as(C.POINTER(C.c_int)),
                                      count.ctypes.data_as(C.POINTER(C.c_int)), nops, _ptr(f), nthreads)
            if rc:
                raise RuntimeError(f"orc_udsc_field failed: {rc}")
            f = f.reshape(V3, nops, 2)
            f = f[..., 0] + 1j * f[..., 1]
            H[:, j] = (dft(L, src[:3], q2max, f, nthreads) if space == "momentum" else f).T
        names += [n for n, _ in tabs]
        blocks.append(H)
    return names, (np.concatenate(blocks) if blocks else np.zeros((0,)))


def tetra_correlator(L, src, totalT, q2max, props, idxs, cols, vals, counts, nthreads=0, space="momentum",
                     views=None):
    """contract_tetraquarks' term sums (lib/kernels/PLEGMA_heavy_light_tetraquarks.cu:26-49)
    on the CPU for one table: props = 4 float64 SoA propagators; idxs, cols
    int [n, 8]; vals [n] complex; counts [nops].  Returns H[k][j][m]."""
    V3 = L[0] * L[1] * L[2]
    s, e, lt = window(L, src, totalT)
    idx = np.ascontiguousarray(np.asarray(idxs, dtype=np.int32).reshape(-1))
    col = np.ascontiguousarray(np.asarray(cols, dtype=np.int32).reshape(-1))
    val = np.ascontiguousarray(np.stack([np.real(vals), np.imag(vals)], -1).astype(np.float64).reshape(-1))
    count = np.asarray(counts, dtype=np.int32)
    first = np.concatenate([[0], np.cumsum(count)[:-1]]).astype(np.int32)
    nops = len(count)
    views = views or [(L[3] * V3, 0, L[3])] * 4
    P = [OrcProp(p.ctypes.data, v[0], v[1], v[2]) for p, v in zip(props, views)]
    ip = C.POINTER(C.c_int)
    H = np.zeros((nops, lt, len(momlist(q2max)) if space == "momentum" else V3), dtype=complex)
    for j in range(lt):
        t = row_time(L, src, totalT, j)
        f = np.zeros(V3 * nops * 2)
        rc = lib().orc_tetra_field(_ivec(L), t, *[C.byref(p) for p in P], idx.ctypes.data_as(ip),
                                   col.ctypes.data_as(ip), _ptr(val), first.ctypes.data_as(ip),
                                   count.ctypes.data_as(ip), nops, _ptr(f), nthreads)
        if rc:
            raise RuntimeError(f"orc_tetra_field failed: {rc}")
        f = f.reshape(V3, nops, 2)
        f = f[..., 0] + 1j * f[..., 1]
        H[:, j] = (dft(L, src[:3], q2max, f, nthreads) if space == "momentum" else f).T
    return H
