"""The C-ABI library builds, loads without a GPU and exports every entry point that
include/ud2.h declares. No compute calls (no GPU here)."""
import os
import re

from unidock_engine import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "ud2.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ud2_[a-z0-9_]+)\s*\(", src)))


def test_header_and_binding_agree():
    decl = declared_functions()
    assert decl == sorted(N.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = N.lib()
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert "gfx950" in N.version()


def test_no_device_is_an_error_code_not_a_crash():
    import ctypes as C
    n = C.c_int32(-1)
    rc = N.lib().ud2_device_count(C.byref(n))
    if rc != 0:
        assert n.value == 0 and N.lib().ud2_last_error()
