"""The C-ABI library loads and exports every function declared in include/*.h
(no compute calls: this runs without a GPU)."""
import ctypes
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", src, flags=re.M):
            if not m.group(0).lstrip().startswith("typedef"):
                names.add(m.group(1))
    return names


def test_library_exports_all_declared_symbols():
    from coffe_amd._lib import LIB_PATH
    assert os.path.exists(LIB_PATH), "engine library not built"
    lib = ctypes.CDLL(LIB_PATH)
    names = _declared()
    assert len(names) >= 20
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_error_strings_host_only():
    from coffe_amd._lib import load_library
    L = load_library()
    for code in range(9):
        assert L.coffe_engine_strerror(code)
    assert L.coffe_engine_last_error() == 0
