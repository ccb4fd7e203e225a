"""CPU checks of the drop-in boundary: libadb.so loads without a GPU and
exports every symbol include/adb_capi.h declares; the struct layouts used by
the Python bindings match the header (sizes computed by the C compiler)."""
import ctypes as C
import os
import re
import subprocess
import tempfile

from adb_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "adb_capi.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(adb_[a-z0-9_]+)\s*\(", src)))


def test_header_matches_binding_list():
    assert set(declared_functions()) == set(capi.EXPORTED_SYMBOLS)


def test_library_exports_all_symbols():
    lib = capi.load_library()
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_struct_sizes_match_header():
    code = r'''
#include <stdio.h>
#include "adb_capi.h"
int main(void) {
  printf("%zu %zu %zu %zu\n", sizeof(adb_scene), sizeof(adb_params), sizeof(adb_step_stats), sizeof(adb_kernel_config));
  return 0;
}'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "s.c")
        exe = os.path.join(d, "s")
        open(c, "w").write(code)
        subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe])
        sizes = list(map(int, subprocess.check_output([exe]).split()))
    assert sizes == [C.sizeof(capi.AdbScene), C.sizeof(capi.AdbParams), C.sizeof(capi.AdbStepStats),
                     C.sizeof(capi.AdbKernelConfig)]


def test_params_default_matches_python():
    lib = capi.load_library()
    p = capi.AdbParams()
    lib.adb_params_default(C.byref(p))
    q = capi.default_params()
    for name, _ in capi.AdbParams._fields_:
        a, b = getattr(p, name), getattr(q, name)
        if name == "gravity":
            assert list(a) == list(b)
        else:
            assert a == b, name


def test_no_device_is_reported_not_crashing():
    # on a CPU-only host the library answers instead of faulting
    n = capi.load_library().adb_device_count()
    assert n >= 0 or n == capi.ADB_E_NO_DEVICE
