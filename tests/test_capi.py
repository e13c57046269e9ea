"""The C-ABI library loads on a CPU-only host and exports every function
declared in include/*.h (no device calls without a GPU; the native loop's
host random streams are checked here, they touch no device)."""
import ctypes
import glob
import os
import re

import pytest

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


@pytest.mark.parametrize("warmup", [0, 1, 623, 700])
def test_mt_stream_continues_numpy_legacy_randomstate(warmup):
    # the native loop's host streams (host code only, no GPU): random_sample()
    # and normal(loc, scale) continue a numpy RandomState bit for bit, across a
    # state regeneration and with a cached Gaussian, and hand the state back
    import ctypes
    import numpy as np
    from mgkmc_amd import _lib
    lib = _lib.load()
    rs = np.random.RandomState(2024)
    rs.random_sample(warmup)
    rs.normal()  # leaves the second polar value cached
    st = _lib.mt_from(rs)
    out = (ctypes.c_double * 1500)()
    assert lib.mgk_mt_fill(ctypes.byref(st), 1, 2.2, 0.45, 1500, out) == 0
    assert np.array_equal(np.array(out[:]), rs.normal(2.2, 0.45, 1500))
    assert lib.mgk_mt_fill(ctypes.byref(st), 0, 0.0, 0.0, 1500, out) == 0
    assert np.array_equal(np.array(out[:]), rs.random_sample(1500))
    other = np.random.RandomState(0)
    _lib.mt_to(other, st)
    assert np.array_equal(other.normal(size=7), rs.normal(size=7))
