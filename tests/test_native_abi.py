"""The C ABI library loads and exports every symbol declared in include/*.h
(no device calls: runs without a GPU)."""
import ctypes
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z_0-9]*\s*\*?\s*(lb_[a-z_0-9]+)\s*\(", src, re.M):
            names.add(m.group(1))
    return names


def test_library_exports_all_declared_symbols():
    from jyulb import _native
    L = _native.lib()
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_library_is_in_tree():
    from jyulb import _native
    assert os.path.commonpath([_native.LIB_PATH, ROOT]) == ROOT
    assert isinstance(_native.lib(), ctypes.CDLL)
