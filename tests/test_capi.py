"""The C-ABI library loads on a CPU-only host and exports every function
declared in include/*.h (no compute calls without a GPU)."""
import ctypes
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b(mgk_[a-z0-9_]+)\s*\(", txt):
            names.add(m.group(1))
    return names


def test_header_symbols_exported():
    from mgkmc_amd import _lib
    assert os.path.exists(_lib.LIB_PATH), "build first: python -m mgkmc_amd.build"
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared()
    assert len(names) >= 5
    for n in sorted(names):
        assert hasattr(lib, n), n
    # the Python binding covers the same set
    assert names == set(_lib.SIGNATURES), names ^ set(_lib.SIGNATURES)


def test_version_without_gpu():
    from mgkmc_amd import _lib
    assert _lib.load().mgk_version() >= 1
