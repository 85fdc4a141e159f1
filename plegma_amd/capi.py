"""ctypes binding of the C ABI in include/plegma_amd.h (libplegma_amd.so).

The library is the product: this module loads it or raises.  There is no CPU
fallback.  torch supplies device memory and streams only.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libplegma_amd.so")

# enums of include/plegma_amd.h
F64, F32 = 0, 1
MODE_F64, MODE_MIXED = 0, 1
SPLIT_BF16X3, SPLIT_BF16X6 = 3, 6
POSITION_SPACE, MOMENTUM_SPACE = 0, 1
MESONS, MESONS16, MESONS_NEW, MESONS_ALL, BARYONS, MESON_MESON, TWOPT = range(7)
G_NUCL = 1 << 0
G_DELTAPP_11, G_DELTAPP_22, G_DELTAPP_33 = 1 << 7, 1 << 8, 1 << 9
G_DELTAPP_ALL = G_DELTAPP_11 | G_DELTAPP_22 | G_DELTAPP_33
BATCH_GRAPH, BATCH_TUNE = 1, 2

STATUS = {0: "PLG_OK", -1: "PLG_ERR_INVALID", -2: "PLG_ERR_NOT_INIT", -3: "PLG_ERR_HIP",
          -4: "PLG_ERR_ALLOC", -5: "PLG_ERR_SITE_SIZE", -6: "PLG_ERR_UNSUPPORTED",
          -7: "PLG_ERR_RANGE"}

EXPORTS = ["plg_init", "plg_finalize", "plg_last_error", "plg_version", "plg_momlist",
           "plg_window", "plg_row_time", "plg_kind_shape", "plg_correlator",
           "plg_correlator_static", "plg_correlator_size", "plg_batch_create", "plg_batch_run", "plg_batch_info",
           "plg_batch_destroy", "plg_batch_profile", "plg_synth_prop", "plg_synth_prop_perturbed", "plg_cast_prop",
           "plg_mlp_create", "plg_mlp_nparams", "plg_mlp_set_params", "plg_mlp_get_params",
           "plg_mlp_forward", "plg_mlp_gradient", "plg_mlp_train", "plg_mlp_destroy", "plg_ft_apply",
           "plg_prop_rotate_physical", "plg_prop_apply_boundaries", "plg_udsc_count", "plg_udsc_product",
           "plg_udsc_op_name", "plg_baryons_udsc_select", "plg_baryons_udsc_size", "plg_baryons_udsc",
           "plg_baryons_udsc_profile", "plg_baryons_table",
           "plg_tetraquark_table", "plg_tetra_count", "plg_tetra_product", "plg_tetra_op_name",
           "plg_tetraquarks_select", "plg_tetraquarks_size", "plg_tetraquarks",
           "plg_tetra_bcud_count", "plg_tetra_bcud_product", "plg_tetra_bcud_op_name",
           "plg_tetraquarks_bcud_select", "plg_tetraquarks_bcud_size", "plg_tetraquarks_bcud"]


class PlgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class plg_prop(C.Structure):
    _fields_ = [("data", C.c_void_p), ("precision", C.c_int), ("t_first", C.c_int),
                ("nt", C.c_int), ("pad_", C.c_int), ("comp_stride", C.c_int64),
                ("plane_exp", C.c_void_p), ("rotate_sign", C.c_int), ("bc_t0", C.c_int)]


class plg_job(C.Structure):
    _fields_ = [("source", C.c_int * 4), ("totalT", C.c_int), ("prop", C.c_int),
                ("prop2", C.c_int), ("out_offset", C.c_int64)]


_lib = None


def lib():
    """Load libplegma_amd.so; raise if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        ip = C.POINTER(C.c_int)
        L.plg_last_error.restype = C.c_char_p
        L.plg_init.argtypes = [ip, C.c_int]
        L.plg_momlist.argtypes = [C.c_int, ip, ip, C.c_int]
        L.plg_window.argtypes = [ip, C.c_int, ip, ip, ip]
        L.plg_row_time.argtypes = [ip, C.c_int, C.c_int, ip]
        L.plg_kind_shape.argtypes = [C.c_int, C.c_int, ip, ip, ip]
        L.plg_correlator.argtypes = [C.c_int, C.c_int, C.POINTER(plg_prop), C.POINTER(plg_prop), ip,
                                     C.c_int, ip, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                     C.c_void_p]
        L.plg_correlator_static.argtypes = L.plg_correlator.argtypes
        L.plg_correlator_size.argtypes = [C.c_int, C.c_int, ip, C.c_int, C.c_int, C.c_int,
                                          C.POINTER(C.c_int64)]
        L.plg_batch_create.argtypes = [C.c_int, C.c_int, C.POINTER(plg_prop), C.c_int,
                                       C.POINTER(plg_job), C.c_int, ip, C.c_int, C.c_int, C.c_int,
                                       C.c_int, C.POINTER(C.c_void_p)]
        L.plg_batch_run.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.plg_batch_info.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64), ip]
        L.plg_batch_destroy.argtypes = [C.c_void_p]
        L.plg_batch_profile.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, ip,
                                        C.POINTER(C.c_double), C.c_char_p]
        L.plg_synth_prop.argtypes = [C.POINTER(plg_prop), C.c_uint64, C.c_uint32, C.c_int,
                                     C.c_double, C.c_void_p]
        L.plg_synth_prop_perturbed.argtypes = [C.POINTER(plg_prop), C.c_uint64, C.c_uint32,
                                               C.c_uint32, C.c_double, C.c_int, C.c_double,
                                               C.c_void_p]
        L.plg_cast_prop.argtypes = [C.POINTER(plg_prop), C.POINTER(plg_prop), C.c_void_p]
        L.plg_ft_apply.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, ip, C.c_int, C.c_int, C.c_int,
                                   C.c_void_p, C.c_int, C.c_void_p]
        fp = C.POINTER(C.c_float)
        pp = C.POINTER(plg_prop)
        L.plg_udsc_count.argtypes = [ip]
        L.plg_udsc_product.argtypes = [C.c_int, C.c_char_p, ip]
        L.plg_udsc_op_name.argtypes = [C.c_int, C.c_int, C.c_char_p, C.c_int]
        L.plg_baryons_udsc_select.argtypes = [pp, pp, pp, pp, C.c_int, C.c_int, ip, C.c_int, ip, ip]
        L.plg_baryons_udsc_size.argtypes = [pp, pp, pp, pp, C.c_int, C.c_int, ip, C.c_int, C.c_int, C.c_int,
                                            C.POINTER(C.c_int64)]
        L.plg_baryons_udsc.argtypes = [pp, pp, pp, pp, C.c_int, C.c_int, ip, C.c_int, ip, C.c_int, C.c_int,
                                       C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        L.plg_baryons_udsc_profile.argtypes = [pp, pp, pp, pp, C.c_int, C.c_int, ip, C.c_int, ip, C.c_int,
                                               C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int, ip,
                                               C.POINTER(C.c_double), C.c_char_p]
        L.plg_baryons_table.argtypes = [pp, pp, pp, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, ip, C.c_int, ip,
                                        C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        L.plg_tetraquark_table.argtypes = [pp, pp, pp, pp, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                           ip, C.c_int, ip, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                           C.c_void_p]
        L.plg_tetra_count.argtypes = [ip]
        L.plg_tetra_product.argtypes = [C.c_int, C.c_char_p, ip]
        L.plg_tetra_op_name.argtypes = [C.c_int, C.c_int, C.c_char_p, C.c_int]
        L.plg_tetraquarks_select.argtypes = [pp, pp, pp, pp, C.c_int, C.c_int, ip, C.c_int, ip, ip]
        L.plg_tetraquarks_size.argtypes = [pp, pp, pp, pp, C.c_int, C.c_int, ip, C.c_int, C.c_int, C.c_int,
                                           C.POINTER(C.c_int64)]
        L.plg_tetraquarks.argtypes = [pp, pp, pp, pp, C.c_int, C.c_int, ip, C.c_int, ip, C.c_int, C.c_int,
                                      C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        L.plg_tetra_bcud_count.argtypes = L.plg_tetra_count.argtypes
        L.plg_tetra_bcud_product.argtypes = L.plg_tetra_product.argtypes
        L.plg_tetra_bcud_op_name.argtypes = L.plg_tetra_op_name.argtypes
        L.plg_tetraquarks_bcud_select.argtypes = L.plg_tetraquarks_select.argtypes
        L.plg_tetraquarks_bcud_size.argtypes = L.plg_tetraquarks_size.argtypes
        L.plg_tetraquarks_bcud.argtypes = L.plg_tetraquarks.argtypes
        L.plg_mlp_create.argtypes = [ip, C.c_int, C.c_int, C.POINTER(C.c_void_p)]
        L.plg_mlp_nparams.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
        L.plg_mlp_set_params.argtypes = [C.c_void_p, fp]
        L.plg_mlp_get_params.argtypes = [C.c_void_p, fp]
        L.plg_mlp_forward.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        L.plg_mlp_gradient.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                       C.POINTER(C.c_double), fp, C.c_void_p]
        L.plg_mlp_train.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_float,
                                    C.POINTER(C.c_double), C.c_void_p]
        L.plg_mlp_destroy.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise PlgError(rc, lib().plg_last_error().decode())
    return rc


def ivec(v):
    return (C.c_int * len(v))(*[int(x) for x in v])


def _stream(s):
    if s is None:
        return None
    return C.c_void_p(s.cuda_stream if hasattr(s, "cuda_stream") else int(s))


# ---------------------------------------------------------------- wrappers
def init(L, device=0):
    check(lib().plg_init(ivec(L), device))


def finalize():
    check(lib().plg_finalize())


def momlist(q2max):
    n = C.c_int()
    check(lib().plg_momlist(q2max, C.byref(n), None, 0))
    buf = (C.c_int * (3 * n.value))()
    check(lib().plg_momlist(q2max, C.byref(n), buf, n.value))
    return [tuple(buf[3 * i:3 * i + 3]) for i in range(n.value)]


def window(source, totalT):
    s, e, lt = C.c_int(), C.c_int(), C.c_int()
    check(lib().plg_window(ivec(source), totalT, C.byref(s), C.byref(e), C.byref(lt)))
    return s.value, e.value, lt.value


def row_time(source, totalT, j):
    t = C.c_int()
    check(lib().plg_row_time(ivec(source), totalT, j, C.byref(t)))
    return t.value


def kind_shape(kind, mask=G_NUCL):
    a, b, c = C.c_int(), C.c_int(), C.c_int()
    check(lib().plg_kind_shape(kind, mask, C.byref(a), C.byref(b), C.byref(c)))
    return a.value, b.value, c.value


def prop_view(tensor, precision, t_first=0, nt=None, V3=None, T=None):
    """plg_prop over a torch tensor holding 144 x (nt*V3) complex entries."""
    nt = T if nt is None else nt
    cs = tensor.numel() // (2 * 144)
    return plg_prop(tensor.data_ptr(), precision, t_first, nt, 0, cs)


def correlator(kind, mask, p1, p2, source, totalT, moms, mode, space, out_prec, out,
               stream=None):
    flat = [int(v) for m in moms for v in m]
    check(lib().plg_correlator(kind, mask, C.byref(p1), C.byref(p2) if p2 is not None else None,
                               ivec(source), totalT, ivec(flat) if flat else None, len(moms), mode,
                               space, out_prec, C.c_void_p(out.data_ptr()), _stream(stream)))


def correlator_static(kind, mask, p1, p2, source, totalT, moms, mode, space, out_prec, out,
                      stream=None):
    """plg_correlator without device allocation (the library's static scratch)"""
    flat = [int(v) for m in moms for v in m]
    check(lib().plg_correlator_static(kind, mask, C.byref(p1), C.byref(p2) if p2 is not None else None,
                                      ivec(source), totalT, ivec(flat) if flat else None, len(moms), mode,
                                      space, out_prec, C.c_void_p(out.data_ptr()), _stream(stream)))


STATIC_SCRATCH_BYTES = 128 << 20


def ft_apply(field, precision, dof, nt, moms, sign, out, out_prec=F64, accumulate=False, stream=None):
    """moms: triples, or a C-contiguous int32 numpy array of shape (nmom, 3),
    which is passed without conversion (a caller applying one list to many
    fields builds it once)"""
    import numpy as np
    if isinstance(moms, np.ndarray) and moms.dtype == np.int32 and moms.ndim == 2 and moms.shape[1] == 3 \
            and moms.flags.c_contiguous:
        mp, n = moms.ctypes.data_as(C.POINTER(C.c_int)), moms.shape[0]
    else:
        flat = [int(v) for m in moms for v in m]
        mp, n = ivec(flat), len(flat) // 3
    check(lib().plg_ft_apply(C.c_void_p(field.data_ptr()), precision, dof, nt, mp, n, sign,
                             int(accumulate), C.c_void_p(out.data_ptr()), out_prec, _stream(stream)))


def udsc_products():
    """[(flavour letters, [dataset names])] of the udsc baryon products, in order"""
    n = C.c_int()
    check(lib().plg_udsc_count(C.byref(n)))
    out = []
    for p in range(n.value):
        fl = C.create_string_buffer(4)
        k = C.c_int()
        check(lib().plg_udsc_product(p, fl, C.byref(k)))
        names = []
        for i in range(k.value):
            b = C.create_string_buffer(64)
            check(lib().plg_udsc_op_name(p, i, b, 64))
            names.append(b.value.decode())
        out.append((fl.value.decode(), names))
    return out


def _ref(v):
    return C.byref(v) if v is not None else None


def udsc_select(up, dn, st, ch, only_st=False, only_ch=False):
    """(product indices, number of datasets) that baryons_udsc computes"""
    prods = (C.c_int * 32)()
    n, nd = C.c_int(), C.c_int()
    check(lib().plg_baryons_udsc_select(_ref(up), _ref(dn), _ref(st), _ref(ch), int(only_st), int(only_ch),
                                        prods, 32, C.byref(n), C.byref(nd)))
    return list(prods[:n.value]), nd.value


def baryons_udsc(up, dn, st, ch, source, totalT, moms, mode, space, out_prec, out, only_st=False, only_ch=False,
                 stream=None):
    """contractBaryonsUDSC: views up/dn/st/ch (None = absent) -> out [dataset][j][m]"""
    flat = [int(v) for m in moms for v in m]
    check(lib().plg_baryons_udsc(_ref(up), _ref(dn), _ref(st), _ref(ch), int(only_st), int(only_ch),
                                 ivec(source), totalT, ivec(flat) if flat else None, len(moms), mode, space,
                                 out_prec, C.c_void_p(out.data_ptr()), _stream(stream)))


def baryons_udsc_profile(up, dn, st, ch, source, totalT, moms, mode, space, out_prec, out, only_st=False,
                         only_ch=False, stream=None, cap=16):
    """one baryons_udsc pass with events between stages: {stage: ms summed over products}"""
    flat = [int(v) for m in moms for v in m]
    n = C.c_int()
    ms = (C.c_double * cap)()
    names = C.create_string_buffer(32 * cap)
    check(lib().plg_baryons_udsc_profile(_ref(up), _ref(dn), _ref(st), _ref(ch), int(only_st), int(only_ch),
                                         ivec(source), totalT, ivec(flat) if flat else None, len(moms), mode, space,
                                         out_prec, C.c_void_p(out.data_ptr()), _stream(stream), cap, C.byref(n), ms,
                                         names))
    raw = names.raw
    return {raw[32 * i:32 * i + 32].split(b"\0")[0].decode(): ms[i] for i in range(min(n.value, cap))}


def baryons_table(p1, p2, p3, idxs, vals, counts, source, totalT, moms, mode, space, out_prec, out, stream=None):
    """table-driven three-propagator contraction (the reference's udsc table
    format): idxs int16 [n, 6], vals float32 [n, 2], counts int16 [nops];
    out [nops][j][m] complex"""
    import numpy as np
    idxs = np.ascontiguousarray(idxs, dtype=np.int16)
    vals = np.ascontiguousarray(vals, dtype=np.float32)
    counts = np.ascontiguousarray(counts, dtype=np.int16)
    flat = [int(v) for m in moms for v in m]
    check(lib().plg_baryons_table(C.byref(p1), C.byref(p2), C.byref(p3), idxs.ctypes.data, vals.ctypes.data,
                                  counts.ctypes.data, len(counts), ivec(source), totalT,
                                  ivec(flat) if flat else None, len(moms), mode, space, out_prec,
                                  C.c_void_p(out.data_ptr()), _stream(stream)))


def tetraquark_table(p1, p2, p3, p4, idxs, cols, vals, counts, source, totalT, moms, mode, space, out_prec, out,
                     stream=None):
    """four-propagator table contraction (the reference's tetraquark table
    format): idxs, cols int16 [n, 8], vals float32 [n, 2], counts int16
    [nops]; out [nops][j][m] complex"""
    import numpy as np
    idxs = np.ascontiguousarray(idxs, dtype=np.int16)
    cols = np.ascontiguousarray(cols, dtype=np.int16)
    vals = np.ascontiguousarray(vals, dtype=np.float32)
    counts = np.ascontiguousarray(counts, dtype=np.int16)
    flat = [int(v) for m in moms for v in m]
    check(lib().plg_tetraquark_table(C.byref(p1), C.byref(p2), C.byref(p3), C.byref(p4), idxs.ctypes.data,
                                     cols.ctypes.data, vals.ctypes.data, counts.ctypes.data, len(counts),
                                     ivec(source), totalT, ivec(flat) if flat else None, len(moms), mode, space,
                                     out_prec, C.c_void_p(out.data_ptr()), _stream(stream)))


_TETRA_FAMILY = {"heavy-light": "", "bcud": "_bcud"}


def _tf(name, family):
    """the C entry point of a tetraquark family: plg_tetra_count ->
    plg_tetra_bcud_count, plg_tetraquarks -> plg_tetraquarks_bcud, ..."""
    sfx = _TETRA_FAMILY[family]
    if name.startswith("plg_tetraquarks"):
        return getattr(lib(), "plg_tetraquarks" + sfx + name[len("plg_tetraquarks"):])
    return getattr(lib(), "plg_tetra" + sfx + name[len("plg_tetra"):])


def tetra_products(family="heavy-light"):
    """[(flavour letters, [dataset names])] of the tetraquark products of a
    family ("heavy-light" or "bcud"), in order"""
    n = C.c_int()
    check(_tf("plg_tetra_count", family)(C.byref(n)))
    out = []
    for p in range(n.value):
        fl = C.create_string_buffer(5)
        k = C.c_int()
        check(_tf("plg_tetra_product", family)(p, fl, C.byref(k)))
        names = []
        for i in range(k.value):
            b = C.create_string_buffer(64)
            check(_tf("plg_tetra_op_name", family)(p, i, b, 64))
            names.append(b.value.decode())
        out.append((fl.value.decode(), names))
    return out


def tetraquarks_select(lt, st, ch, bt, only_st=False, only_ch=False, family="heavy-light"):
    """(product indices, number of datasets) that tetraquarks computes"""
    prods = (C.c_int * 32)()
    n, nd = C.c_int(), C.c_int()
    check(_tf("plg_tetraquarks_select", family)(_ref(lt), _ref(st), _ref(ch), _ref(bt), int(only_st),
                                                int(only_ch), prods, 32, C.byref(n), C.byref(nd)))
    return list(prods[:n.value]), nd.value


def tetraquarks_size(lt, st, ch, bt, source, totalT, nmom, space=MOMENTUM_SPACE, only_st=False, only_ch=False,
                     family="heavy-light"):
    n = C.c_int64()
    check(_tf("plg_tetraquarks_size", family)(_ref(lt), _ref(st), _ref(ch), _ref(bt), int(only_st), int(only_ch),
                                              ivec(source), totalT, nmom, space, C.byref(n)))
    return n.value


def tetraquarks(lt, st, ch, bt, source, totalT, moms, mode, space, out_prec, out, only_st=False, only_ch=False,
                stream=None, family="heavy-light"):
    """contractTetraquarks (family "heavy-light") or contractTetraquarksBCUD
    (family "bcud"): views lt/st/ch/bt (None = absent) -> out [dataset][j][m]"""
    flat = [int(v) for m in moms for v in m]
    check(_tf("plg_tetraquarks", family)(_ref(lt), _ref(st), _ref(ch), _ref(bt), int(only_st), int(only_ch),
                                         ivec(source), totalT, ivec(flat) if flat else None, len(moms), mode,
                                         space, out_prec, C.c_void_p(out.data_ptr()), _stream(stream)))


def baryons_udsc_size(up, dn, st, ch, source, totalT, nmom, space=MOMENTUM_SPACE, only_st=False, only_ch=False):
    n = C.c_int64()
    check(lib().plg_baryons_udsc_size(_ref(up), _ref(dn), _ref(st), _ref(ch), int(only_st), int(only_ch),
                                      ivec(source), totalT, nmom, space, C.byref(n)))
    return n.value


def correlator_size(kind, mask, source, totalT, nmom, space=MOMENTUM_SPACE):
    n = C.c_int64()
    check(lib().plg_correlator_size(kind, mask, ivec(source), totalT, nmom, space, C.byref(n)))
    return n.value


class Batch:
    """plg_batch_* wrapper: many jobs in one plan.  A job is
    (source, totalT, prop, out_offset[, prop2]) with indices into props."""

    def __init__(self, kind, mask, props, jobs, moms, mode, split=SPLIT_BF16X6, flags=0):
        P = (plg_prop * len(props))(*props)
        J = (plg_job * len(jobs))()
        for i, job in enumerate(jobs):
            src, totalT, prop, off = job[:4]
            J[i].source = (C.c_int * 4)(*src)
            J[i].totalT = totalT
            J[i].prop = prop
            J[i].prop2 = job[4] if len(job) > 4 else -1
            J[i].out_offset = off
        flat = [int(v) for m in moms for v in m]
        h = C.c_void_p()
        check(lib().plg_batch_create(kind, mask, P, len(props), J, len(jobs), ivec(flat),
                                     len(moms), mode, split, flags, C.byref(h)))
        self.h = h
        self._keep = (P, J)

    def info(self):
        a, b, c = C.c_int64(), C.c_int64(), C.c_int()
        check(lib().plg_batch_info(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def profile(self, out, stream=None, cap=16):
        """[(stage name, milliseconds)] of one eager pass with events between stages"""
        n = C.c_int()
        ms = (C.c_double * cap)()
        names = C.create_string_buffer(32 * cap)
        check(lib().plg_batch_profile(self.h, C.c_void_p(out.data_ptr()), _stream(stream), cap,
                                      C.byref(n), ms, names))
        raw = names.raw
        return [(raw[32 * i:32 * i + 32].split(b"\0")[0].decode(), ms[i]) for i in range(n.value)]

    def run(self, out, stream=None):
        check(lib().plg_batch_run(self.h, C.c_void_p(out.data_ptr()), _stream(stream)))

    def close(self):
        if self.h:
            lib().plg_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def synth_prop(view, seed, stream_id, t0, lam, stream=None):
    check(lib().plg_synth_prop(C.byref(view), seed, stream_id, t0, lam, _stream(stream)))


def synth_prop_perturbed(view, seed, base_stream, pert_stream, eps, t0, lam, stream=None):
    check(lib().plg_synth_prop_perturbed(C.byref(view), seed, base_stream, pert_stream, eps, t0,
                                         lam, _stream(stream)))


def cast_prop(dst, src, stream=None):
    check(lib().plg_cast_prop(C.byref(dst), C.byref(src), _stream(stream)))


def prop_rotate_physical(view, sign, stream=None):
    check(lib().plg_prop_rotate_physical(C.byref(view), sign, _stream(stream)))


def prop_apply_boundaries(view, t0, stream=None):
    check(lib().plg_prop_apply_boundaries(C.byref(view), t0, _stream(stream)))


class Mlp:
    """plg_mlp_* wrapper.  Inputs/targets/outputs are device tensors (torch,
    float32, contiguous [batch][features]); parameters travel as numpy float32."""

    def __init__(self, dims, max_batch):
        self.dims = [int(d) for d in dims]
        h = C.c_void_p()
        check(lib().plg_mlp_create(ivec(self.dims), len(self.dims) - 1, max_batch, C.byref(h)))
        self.h = h
        n = C.c_int64()
        check(lib().plg_mlp_nparams(self.h, C.byref(n)))
        self.nparams = n.value

    def set_params(self, p):
        import numpy as np
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparams
        check(lib().plg_mlp_set_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float))))

    def get_params(self):
        import numpy as np
        p = np.empty(self.nparams, dtype=np.float32)
        check(lib().plg_mlp_get_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float))))
        return p

    def forward(self, x, y, stream=None):
        check(lib().plg_mlp_forward(self.h, C.c_void_p(x.data_ptr()), x.shape[0],
                                    C.c_void_p(y.data_ptr()), _stream(stream)))

    def gradient(self, x, t, stream=None):
        import numpy as np
        g = np.empty(self.nparams, dtype=np.float32)
        loss = C.c_double()
        check(lib().plg_mlp_gradient(self.h, C.c_void_p(x.data_ptr()), C.c_void_p(t.data_ptr()),
                                     x.shape[0], C.byref(loss),
                                     g.ctypes.data_as(C.POINTER(C.c_float)), _stream(stream)))
        return loss.value, g

    def train(self, x, t, steps, lr, stream=None):
        loss = C.c_double()
        check(lib().plg_mlp_train(self.h, C.c_void_p(x.data_ptr()), C.c_void_p(t.data_ptr()),
                                  x.shape[0], steps, lr, C.byref(loss), _stream(stream)))
        return loss.value

    def close(self):
        if self.h:
            lib().plg_mlp_destroy(self.h)
            self.h = None
