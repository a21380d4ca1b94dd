"""The C-ABI library loads without a GPU and exports every entry point include/vca.h declares."""
import ctypes
import os
import re

from volcano_ca_mc_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "vca.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*int\s+(vca_\w+)\s*\(", txt, flags=re.M)))


def test_header_and_binding_agree():
    assert declared() == sorted(native.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(native.LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), name


def test_no_device_is_reported_cleanly_or_device_is_gfx950():
    # Without a GPU the context call must fail with an error code and a message, never crash.
    lib = native.lib()
    h = ctypes.c_void_p()
    rc = lib.vca_ctx_create(0, None, ctypes.byref(h))
    if rc == 0:
        lib.vca_ctx_destroy(h)
    else:
        buf = ctypes.create_string_buffer(256)
        lib.vca_last_error(buf, 256)
        assert buf.value
