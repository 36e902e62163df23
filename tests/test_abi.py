"""The C ABI library (libspock_amd.so) loads on a machine without a GPU and exports every entry
point include/spock_amd.h declares; the Python binding covers all of them.  No compute calls."""
import ctypes as C
import os
import re

import pytest

from spock_amd import runtime as rt

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "spock_amd.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(spk_\w+)\s*\(", text)))


def test_header_declares_the_binding_exports():
    assert set(declared()) == set(rt.EXPORTS)


@pytest.mark.skipif(not os.path.exists(rt.LIB_PATH), reason="library not built (run __graft_entry__.build())")
def test_library_exports_every_declared_symbol():
    lib = C.CDLL(rt.LIB_PATH)
    missing = [s for s in declared() if not hasattr(lib, s)]
    assert not missing


@pytest.mark.skipif(not os.path.exists(rt.LIB_PATH), reason="library not built (run __graft_entry__.build())")
def test_default_options_are_the_reference_builder_defaults():
    """src/cache.cuh:1154-1210: maxInner 8, Anderson memory 3, error check every 25 iterations."""
    o = rt.options()
    assert (o.max_inner, o.anderson_buffer, o.check_period, o.method) == (8, 3, 25, rt.SPOCK)
    assert o.tol_abs == pytest.approx(1e-3) and o.tol_rel == 0.0 and o.max_iters == 0
