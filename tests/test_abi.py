"""The C-ABI library loads and exports every entry point that
include/vhlle_gpu.h declares (no compute calls: runs without a GPU)."""
import os
import re

from vhlle_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = []
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", f)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names += re.findall(r"\b(?:int|long|void)\s+(vh_\w+)\s*\(", txt)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    L = abi.lib()
    names = declared()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(abi.exported_symbols()) == names


def test_no_device_is_an_error_not_a_fallback():
    """Without a GPU the library refuses to create a context (no CPU path)."""
    import ctypes
    L = abi.lib()
    n = ctypes.c_int(0)
    g = abi.Grid(4, 4, 1, -1, 1, -1, 1, -1, 1)
    h = ctypes.c_void_p()
    rc = L.vh_create(ctypes.byref(g), ctypes.byref(abi.Eos(0)), ctypes.byref(abi.make_trc()),
                     0, ctypes.byref(h))
    if rc == abi.VH_OK:  # a GPU is present: clean up
        L.vh_destroy(h)
    else:
        assert rc == abi.VH_ERR_NODEVICE
