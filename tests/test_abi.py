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


def _struct_fields(name):
    """(field, C type) of `typedef struct name {...} name;` in the header, arrays as 'type[n]'."""
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (name, name), text, flags=re.S).group(1)
    out = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        m = re.match(r"(double|long|int)\s+(.*)", decl)
        ctype, names = m.group(1), m.group(2)
        for nm in names.split(","):
            nm = nm.strip()
            arr = re.match(r"(\w+)\[(\d+)\]", nm)
            out.append((arr.group(1), f"{ctype}[{arr.group(2)}]") if arr else (nm, ctype))
    return out


_CTYPES = {C.c_double: "double", C.c_long: "long", C.c_int: "int"}


def _binding_fields(cls):
    out = []
    for nm, t in cls._fields_:
        if hasattr(t, "_length_"):
            out.append((nm, f"{_CTYPES[t._type_]}[{t._length_}]"))
        else:
            out.append((nm, _CTYPES[t]))
    return out


@pytest.mark.parametrize("cname,cls", [("spk_options", rt.Options), ("spk_result", rt.Result)])
def test_binding_structs_match_the_header(cname, cls):
    """The ctypes mirrors of spk_options / spk_result have the header's fields, order and types."""
    assert _binding_fields(cls) == _struct_fields(cname)


@pytest.mark.skipif(os.path.exists("/dev/kfd") or not os.path.exists(rt.LIB_PATH),
                    reason="needs the built library on a machine without a GPU")
def test_no_device_fails_loudly():
    """Without a HIP device the library reports an error instead of falling back to the CPU."""
    with pytest.raises(rt.SpockError):
        rt.Context(0)
