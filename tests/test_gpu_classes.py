"""The C++ class layer (include/plegma/: PLEGMA_Propagator, PLEGMA_FT,
PLEGMA_Correlator over the C ABI), driven by tests/cpp/class_driver.cpp as a
PLEGMA driver would use it, against the CPU oracle on the same seeded
propagators: H_elem layout and values, and the ASCII file."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from plegma_amd import build

import gpu_util as G

pytestmark = pytest.mark.gpu

L = [4, 4, 4, 8]
SRC = [1, 2, 3, 5]


@pytest.fixture(scope="module")
def outputs(tmp_path_factory):
    G.torch_cuda()
    d = tmp_path_factory.mktemp("classes")
    for prec in ("f64", "f32"):
        print(f"[gpu] class driver {prec}", flush=True)
        r = subprocess.run([build.DRIVER, str(d), prec], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
    u = O.synth_prop(L, 1234, 1, 0, 0.5)
    dn = O.synth_prop(L, 1234, 2, 0, 0.5)
    return d, u, dn


def _load(d, name, prec):
    a = np.fromfile(os.path.join(d, f"{name}_{prec}.bin"), dtype=np.float64 if prec == "f64" else np.float32)
    a = a.astype(np.float64)
    return a[0::2] + 1j * a[1::2]


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 1e-6)])
@pytest.mark.parametrize("name,kind,q2,totalT,mask", [
    ("mesons", O.MESONS, 4, 6, O.G_NUCL),
    ("baryons", O.BARYONS, 4, 6, O.G_NUCL | sum(O.G_DELTA32)),
    ("mesons_all", O.MESONS_ALL, 1, 8, O.G_NUCL),
    ("mesmes", O.MESON_MESON, 4, 6, O.G_NUCL)])
def test_class_results(outputs, prec, tol, name, kind, q2, totalT, mask):
    d, u, dn = outputs
    got = _load(d, name, prec)
    ref = O.correlator(kind, L, SRC, totalT, q2, u, dn, mask=mask)
    assert got.size == ref.size
    got = got.reshape(ref.shape)
    assert G.rel_err(got, ref) < tol


def test_ascii_file(outputs):
    """lines 'it nx ny nz (re imI) x site size' ordered by time from the source
    (lib/PLEGMA_Correlator.cu:786-797); totalT = T so every time is present"""
    d, u, dn = outputs
    path = os.path.join(d, "mesons_all_f64_sx01sy02sz03st05.dat")
    lines = open(path).read().splitlines()
    moms = O.momlist(1)
    assert len(lines) == 8 * len(moms)
    ref = O.correlator(O.MESONS_ALL, L, SRC, 8, 1, u, dn)          # [g][d][j][m][c]
    for ln in (lines[0], lines[len(moms) + 3], lines[-1]):
        tok = ln.split()
        it, n = int(tok[0]), [int(x) for x in tok[1:4]]
        m = [tuple(v) for v in moms.tolist()].index(tuple(n))
        t = (it + SRC[3]) % 8
        j = [O.row_time(L, SRC, 8, r) for r in range(8)].index(t)
        vals = [complex(float(a), float(b[:-1])) for a, b in zip(tok[4::2], tok[5::2])]
        np.testing.assert_allclose(vals, ref[0, 0, j, m], rtol=1e-5, atol=1e-6 * np.abs(ref).max())


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 1e-6)])
def test_ft_apply_class(outputs, prec, tol):
    """PLEGMA_FT<Float>::apply on a propagator: H_elem[t][144][mom].  The FP32
    propagator of the driver stores each time slice times 2^-e with
    e = floor(log2 max |entry|) of the slice (plg_cast_prop); a raw field
    transform sees those stored values."""
    import closed_forms as cf
    d, u, dn = outputs
    got = _load(d, "ft", prec)
    moms = O.momlist(2)
    V3 = 64
    got = got.reshape(8, 144, len(moms))
    U = (u[0::2] + 1j * u[1::2]).reshape(144, 8, V3)
    if prec == "f32":
        amax = np.maximum(np.abs(U.real), np.abs(U.imag)).max(axis=(0, 2))    # per time slice
        e = np.floor(np.log2(amax)).astype(int)
        U = (U * np.ldexp(1.0, -e)[None, :, None]).astype(np.complex64).astype(np.complex128)
    for t in (0, 5):
        ref = cf.dft(U[:, t, :].T, L, [0, 0, 0], moms)       # [mom][144]
        assert G.rel_err(got[t].T, ref, axis=0) < tol


def _descr(body):
    return "shape: " + body + "/re-im"


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 1e-6)])
def test_hdf5_record_file(outputs, prec, tol):
    """writeHDF5 / writeFile(.h5): the objects of the reference's HDF5 file
    (lib/PLEGMA_Correlator.cu:809-946) in the record file: paths
    /sxXXsyXXszXXstXX/<group>/<dataset>, mvec, shape (totalT, Nmom, shape, 2)
    with time counted from the source, the description attribute, and the
    skip of datasets already present."""
    from plegma_amd import h5twin
    d, u, dn = outputs
    ds = h5twin.read(os.path.join(d, f"corr_{prec}.h5.twin"))
    top = "/sx01sy02sz03st05"
    mdesc = "pseudoscalar, scalar, g5g1, g5g2, g5g3, g5g4, g1, g2, g3, g4"
    bdesc = "1,g1,g2,g3,g4,g5,g5g1,g5g2,g5g3,g5g4,s12,s13,s23,s41,s42,s43"
    bgroups = ["nucl_nucl", "deltapp_deltamm_11", "deltapp_deltamm_22", "deltapp_deltamm_33"]
    want = [f"{top}/mesons/mvec", f"{top}/mesons/twop_meson_1", f"{top}/mesons/twop_meson_2"]
    for g in bgroups:
        want += [f"{top}/baryons/{g}/mvec", f"{top}/baryons/{g}/twop_baryon_1", f"{top}/baryons/{g}/twop_baryon_2"]
    want += [f"{top}/mesons/twop_meson"]
    assert list(ds) == want
    dt = np.float64 if prec == "f64" else np.float32
    moms = O.momlist(4)
    np.testing.assert_array_equal(ds[f"{top}/mesons/mvec"].data, moms)   # first writer's list kept

    def check(path, ref_block, totalT, descr):
        obj = ds[path]
        assert obj.data.dtype == dt
        assert obj.attrs == {"description": descr}
        got = obj.data.astype(np.float64)
        got = got[..., 0] + 1j * got[..., 1]
        assert got.shape[0] == totalT and got.shape[1:] == ref_block.shape[1:]
        for j in range(totalT):
            tau = (O.row_time(L, SRC, totalT, j) - SRC[3]) % L[3]
            assert G.rel_err(got[tau], ref_block[j]) < tol

    ref = O.correlator(O.MESONS, L, SRC, 6, 4, u, dn)                    # [g][d][j][m][c]
    for k in range(2):
        check(f"{top}/mesons/twop_meson_{k + 1}", ref[0, k], 6, _descr("/time/moms/" + mdesc))
    ref = O.correlator(O.BARYONS, L, SRC, 6, 4, u, dn, mask=O.G_NUCL | sum(O.G_DELTA32))
    for gi, g in enumerate(bgroups):
        np.testing.assert_array_equal(ds[f"{top}/baryons/{g}/mvec"].data, moms)
        for k in range(2):
            check(f"{top}/baryons/{g}/twop_baryon_{k + 1}", ref[gi, k], 6, _descr("/time/moms/" + bdesc))
    ref = O.correlator(O.MESONS_ALL, L, SRC, 8, 1, u, dn)
    alld = "pseudoscalar, scalar" + "".join(f", ({v}g{i},{v}g{j})" for v in ("g5", "")
                                            for i in range(1, 5) for j in range(1, 5))
    check(f"{top}/mesons/twop_meson", ref[0, 0], 8, _descr("/time/moms/" + alld))


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 1e-6)])
def test_hdf5_position_space(outputs, prec, tol):
    """position space: (totalT, Lz, Ly, Lx, 10, 2), every axis counted from
    the source with periodic wrap (lib/PLEGMA_Correlator.cu:825-832)"""
    from plegma_amd import h5twin
    d, u, dn = outputs
    ds = h5twin.read(os.path.join(d, f"pos_{prec}.h5.twin"))
    path = "/sx01sy02sz03st05/mesons/twop_meson"
    assert list(ds) == [path]
    mdesc = "pseudoscalar, scalar, g5g1, g5g2, g5g3, g5g4, g1, g2, g3, g4"
    assert ds[path].attrs == {"description": _descr("/t/z/y/x/" + mdesc)}
    got = ds[path].data.astype(np.float64)
    got = got[..., 0] + 1j * got[..., 1]
    assert got.shape == (4, L[2], L[1], L[0], 10)
    # the raw H_elem of the same correlator: rows j over x lexicographic
    H = _load(d, "pos_mesons", prec).reshape(4, L[2], L[1], L[0], 10)
    for j in range(4):
        t = O.row_time(L, SRC, 4, j)
        tau = (t - SRC[3]) % L[3]
        ref = O.site_correlator(O.MESONS_NEW, L, t, u, dn).reshape(L[2], L[1], L[0], 10)
        assert G.rel_err(H[j], ref, axis=(0, 1, 2)) < tol
        sh = np.roll(ref, shift=(-SRC[2], -SRC[1], -SRC[0]), axis=(0, 1, 2))
        assert G.rel_err(got[tau], sh, axis=(0, 1, 2)) < tol


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 1e-6)])
def test_udsc_class(outputs, prec, tol):
    """contractBaryonsUDSC(u, d, s, absent charm): datasets of the u/d/s
    products, H_elem [dataset][t][mom], against the oracle"""
    from plegma_amd import gen_udsc
    d, u, dn = outputs
    s = O.synth_prop(L, 1234, 3, 0, 0.8)
    names, ref = O.udsc_correlator(L, SRC, 6, 2, {"u": u, "d": dn, "s": s}, gen_udsc.products())
    got_names = open(os.path.join(d, f"udsc_{prec}.names")).read().split()
    assert got_names == names
    got = _load(d, "udsc", prec).reshape(ref.shape)
    assert G.rel_err(got, ref, axis=-1) < tol


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 1e-6)])
def test_udsc_hdf5(outputs, prec, tol):
    """udsc datasets in the HDF5 record file: no groups and no shape, so every
    operator is /sxXXsyXXszXXstXX/<baryon>/<operator> of shape (totalT, Nmom, 2)
    and mvec sits at the source level"""
    from plegma_amd import gen_udsc, h5twin
    d, u, dn = outputs
    s = O.synth_prop(L, 1234, 3, 0, 0.8)
    names, ref = O.udsc_correlator(L, SRC, 6, 2, {"u": u, "d": dn, "s": s}, gen_udsc.products())
    ds = h5twin.read(os.path.join(d, f"udsc_{prec}.h5.twin"))
    top = "/sx01sy02sz03st05"
    assert list(ds) == [f"{top}/mvec"] + [f"{top}/{n}" for n in names]
    np.testing.assert_array_equal(ds[f"{top}/mvec"].data, O.momlist(2))
    for k, n in enumerate(names):
        obj = ds[f"{top}/{n}"]
        assert obj.attrs == {"description": "shape: /time/moms/re-im"}
        got = obj.data.astype(np.float64)
        got = got[..., 0] + 1j * got[..., 1]
        for j in range(6):
            tau = (O.row_time(L, SRC, 6, j) - SRC[3]) % L[3]
            assert G.rel_err(got[tau][:, None], ref[k, j][:, None], axis=0) < tol


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 1e-6)])
def test_tetraquarks_class(outputs, prec, tol):
    """contractTetraquarks(LT = u, ST = s, CH absent, BT = b) into a double
    correlator: the datasets of "udbb" and "usbb" (reference order, names),
    H_elem [dataset][t][mom], against the oracle's 81-term sums over the
    generated tables"""
    from plegma_amd import gen_tetra
    d, u, dn = outputs
    s = O.synth_prop(L, 1234, 3, 0, 0.8)
    b = O.synth_prop(L, 1234, 4, 0, 1.2)
    fl = {"u": u, "d": u, "s": s, "b": b}
    names, refs = [], []
    for p, ops in gen_tetra.tables():
        idx = np.concatenate([o[1] for o in ops])
        col = np.concatenate([o[2] for o in ops])
        val = np.concatenate([o[3] for o in ops]).astype(np.complex128)
        refs.append(O.tetra_correlator(L, SRC, 6, 2, [fl[f] for f in p], idx, col, val, [len(o[3]) for o in ops]))
        names += [o[0] for o in ops]
    ref = np.concatenate(refs)
    assert open(os.path.join(d, f"tetra_{prec}.names")).read().split() == names
    a = np.fromfile(os.path.join(d, f"tetra_{prec}.bin"), dtype=np.float64)
    got = (a[0::2] + 1j * a[1::2]).reshape(ref.shape)
    assert G.rel_err(got, ref, axis=-1) < tol


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 1e-6)])
def test_tetraquarks_bcud_class(outputs, prec, tol):
    """contractTetraquarksBCUD(LT = u, ST absent, CH = c, BT = b) into a double
    correlator: the 57 datasets of "udcb" (reference order, names), H_elem
    [dataset][t][mom], against the oracle's 81-term sums over the generated
    bcud tables (terms pairing the colours of P1 with P4 and P2 with P3 run
    as a second launch)"""
    from plegma_amd import gen_tetra
    d, u, dn = outputs
    c = O.synth_prop(L, 1234, 5, 0, 1.0)
    b = O.synth_prop(L, 1234, 4, 0, 1.2)
    fl = {"u": u, "d": u, "c": c, "b": b}
    (p, ops), = gen_tetra.bcud_tables()
    idx = np.concatenate([o[1] for o in ops])
    col = np.concatenate([o[2] for o in ops])
    val = np.concatenate([o[3] for o in ops]).astype(np.complex128)
    ref = O.tetra_correlator(L, SRC, 6, 2, [fl[f] for f in p], idx, col, val, [len(o[3]) for o in ops])
    assert open(os.path.join(d, f"bcud_{prec}.names")).read().split() == [o[0] for o in ops]
    a = np.fromfile(os.path.join(d, f"bcud_{prec}.bin"), dtype=np.float64)
    got = (a[0::2] + 1j * a[1::2]).reshape(ref.shape)
    assert G.rel_err(got, ref, axis=-1) < tol
