"""Build recipe of libplegma_amd.so (hipcc, gfx950, in-tree).

Compiles plegma_amd/csrc/*.hip into plegma_amd/libplegma_amd.so.  Objects go
to plegma_amd/_build/, as does class_driver, the C++ program that drives the
class layer of include/plegma/ (its tests run it).  Called by
__graft_entry__.build(); runnable directly:
    python -m plegma_amd.build
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libplegma_amd.so")
SOURCES = ["plg_api.hip", "plg_contract.hip", "plg_baryon.hip", "plg_project.hip", "plg_project_bf16.hip", "plg_project_sep.hip",
           "plg_synth.hip", "plg_mesmes.hip", "plg_mlp.hip", "plg_prep.hip", "plg_udsc.hip",
           "plg_tetra.hip"]
# generated from the udsc baryon and the heavy-light tetraquark interpolating
# fields by plegma_amd/gen_udsc.py and plegma_amd/gen_tetra.py
UDSC_TABLES = os.path.join(CSRC, "plg_udsc_tables.h")
TETRA_TABLES = os.path.join(CSRC, "plg_tetra_tables.h")
ARCH = "gfx950"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
    # no packed FP32 arithmetic in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-variable", "-Wno-pass-failed",
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, src.replace(".hip", ".o"))
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    from plegma_amd import gen_tetra, gen_udsc
    gen_udsc.write_header(UDSC_TABLES)
    gen_tetra.write_header(TETRA_TABLES)
    with cf.ThreadPoolExecutor(max_workers=min(4, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    build_driver()
    if verbose:
        print("built", LIB)
    return LIB


DRIVER = os.path.join(BUILD, "class_driver")


def build_driver() -> str:
    """host C++ (g++) over the class headers, linked to the library in-tree"""
    src = os.path.join(ROOT, "tests", "cpp", "class_driver.cpp")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"), "-I/opt/rocm/include",
           src, "-o", DRIVER + ".tmp", "-L" + HERE, "-lplegma_amd", "-L/opt/rocm/lib", "-lamdhip64",
           "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"class driver build failed:\n{r.stderr}")
    os.replace(DRIVER + ".tmp", DRIVER)
    return DRIVER


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
