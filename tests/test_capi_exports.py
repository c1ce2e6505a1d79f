"""The C-ABI library loads and exports every entry point include/skirt7_amd.h declares (CPU only: no
compute calls)."""
import ctypes
import os
import re

from skirt7_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "skirt7_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(sg_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    names = declared()
    assert len(names) >= 25
    L = ctypes.CDLL(capi.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    assert set(declared()) <= set(capi.EXPORTED)


def test_errors_are_reported_without_a_device_call():
    # a Sersic index outside [0.5, 10] is rejected on the host with a message
    import numpy as np
    s, M = np.zeros(101), np.zeros(101)
    st = capi.lib().sg_sersic_table(0.1, capi.dptr(s), capi.dptr(M))
    assert st != 0 and b"Sersic" in capi.lib().sg_last_error()
