"""CPU-side checks of the C ABI: the library loads, exports every function
include/plegma_amd.h declares, and the entry points that need no device
(momentum list, kind shapes) agree with the oracle.  No kernel is launched."""
import os
import re

import numpy as np
import pytest

from oracle import oracle as O
from plegma_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "plegma_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(plg_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_all_declared_symbols():
    lib = capi.lib()
    names = declared()
    assert names, "no declarations parsed"
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(capi.EXPORTS) == names


def test_momlist_without_device():
    for q in (0, 2, 4, 9, 16):
        got = np.array(capi.momlist(q)).reshape(-1, 3)
        assert np.array_equal(got, O.momlist(q))


def test_negative_q2_is_an_error():
    with pytest.raises(capi.PlgError) as e:
        capi.momlist(-3)
    assert e.value.code == -1


@pytest.mark.parametrize("kind,okind,mask", [
    (capi.MESONS, O.MESONS, 0), (capi.MESONS16, O.MESONS16, 0),
    (capi.MESONS_NEW, O.MESONS_NEW, 0), (capi.MESONS_ALL, O.MESONS_ALL, 0),
    (capi.BARYONS, O.BARYONS, capi.G_NUCL | capi.G_DELTAPP_ALL),
    (capi.MESON_MESON, O.MESON_MESON, 0)])
def test_kind_shapes(kind, okind, mask):
    assert capi.kind_shape(kind, mask) == O.kind_shape(okind, mask)


def test_unsupported_groups_rejected():
    with pytest.raises(capi.PlgError) as e:
        capi.kind_shape(capi.BARYONS, 1 << 1)   # nucl_nucl2: not on this path
    assert e.value.code == -6


def test_calls_before_init_fail_cleanly():
    with pytest.raises(capi.PlgError) as e:
        capi.window([0, 0, 0, 0], 4)
    assert e.value.code == -2


def test_class_headers_compile_and_driver_is_built():
    """the C++ class layer is plain host C++17 over the C ABI"""
    import subprocess
    from plegma_amd import build
    src = os.path.join(ROOT, "tests", "cpp", "class_driver.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror",
                        "-I" + os.path.join(ROOT, "include"), "-I/opt/rocm/include", src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert os.access(build.DRIVER, os.X_OK)


def test_c_header_is_plain_c99():
    """the C ABI header compiles as C99 (the cgo / ctypes side of a binding)"""
    import subprocess
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "t.c")
        with open(src, "w") as f:
            f.write('#include "plegma_amd.h"\nint main(void) { return plg_version() > 0 ? 0 : 1; }\n')
        r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror", "-fsyntax-only",
                            "-I" + os.path.join(ROOT, "include"), src], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


def test_udsc_and_table_calls_before_init_fail_cleanly():
    import numpy as np
    with pytest.raises(capi.PlgError) as e:
        capi.baryons_udsc_size(None, None, None, None, [0, 0, 0, 0], 4, 1)
    assert e.value.code == -2
    v = capi.plg_prop()
    with pytest.raises(capi.PlgError) as e:
        capi.baryons_table(v, v, v, np.zeros((1, 6)), np.ones((1, 2)), np.ones(1), [0, 0, 0, 0], 4, [[0, 0, 0]],
                           capi.MODE_F64, capi.MOMENTUM_SPACE, capi.F64, _FakeTensor())
    assert e.value.code == -2


class _FakeTensor:
    def data_ptr(self):
        return 0
