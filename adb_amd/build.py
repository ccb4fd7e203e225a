"""Build libadb.so (HIP, gfx950) in-tree with hipcc.

Every HIP source is compiled with -ffp-contract=off (FP64 predicates must be
separately rounded operations) and with packed FP32 instructions disabled.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libadb.so")
SOURCES = ["contact.hip", "cull.hip", "assemble.hip", "pcg.hip", "energy.hip", "ccd.hip", "hash.hip", "world.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result"]


def _compile(src):
    obj = os.path.join(BUILD, src.replace(".hip", ".o"))
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr[-6000:]))
    return obj


def build_pybind(verbose=True):
    """The reference's `abdtk` module as a pybind11 extension over libadb.so,
    in adb_amd/pyext/ (import it with that folder on sys.path: `import abdtk`).
    Host code only (g++); the kernels it calls are libadb.so's."""
    import sysconfig
    import pybind11
    out_dir = os.path.join(HERE, "pyext")
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "abdtk" + sysconfig.get_config_var("EXT_SUFFIX"))
    cmd = ["g++", "-O2", "-shared", "-fPIC", "-std=c++17", "-I", pybind11.get_include(),
           "-I", sysconfig.get_paths()["include"], os.path.join(CSRC, "abdtk_pybind.cpp"),
           "-L", HERE, "-ladb", "-Wl,-rpath,$ORIGIN/..", "-o", out + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("pybind11 module build failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-4000:]))
    os.replace(out + ".tmp", out)
    if verbose:
        print("built", out)
    return out


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    # --no-undefined: a symbol left undefined fails here, not at load time
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-Wl,--no-undefined", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s" % r.stderr[-4000:])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_pybind(verbose)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
