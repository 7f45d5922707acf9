"""CPU checks of the C ABI: the library loads without a GPU, exports every
entry point include/titan_hip.h declares, and rejects bad arguments before
touching a device."""
import ctypes as C
import os
import re

import pytest

from titan_amd import binding as B

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "titan_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(thx?_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_abi():
    names = _declared()
    for n in B.EXPORTS:
        assert n in names, n


def test_library_exports_every_declared_symbol():
    lib = B.load()
    for n in _declared():
        assert hasattr(lib, n), "missing export %s" % n


def test_bad_arguments_are_rejected_without_a_device():
    lib = B.load()
    h = C.c_void_p()
    # NULL out pointer / sizes out of range: rejected before any HIP call
    assert lib.th_create(1, 9, 0, None, None, None, None, None, None) == -1
    assert lib.th_create(0, 9, 0, None, None, None, None, None, C.byref(h)) == -1
    assert lib.th_create(1, 10, 0, None, None, None, None, None, C.byref(h)) == -1
    assert b"th_create" in lib.th_last_error()
    assert lib.th_set_imag_points(None, 1, None, None, None) == -1
    assert lib.th_destroy(None) == 0


def test_limits():
    lim = B.limits()
    assert lim["max_nOrb"] == 9 and lim["max_d"] >= 54


def test_product_has_no_cpu_fallback():
    """The product package never imports the oracle or numpy linear algebra
    for its results: a missing library raises."""
    import importlib
    src = open(os.path.join(ROOT, "titan_amd", "binding.py")).read()
    assert "oracle" not in src
    for mod in ("titan_amd.system", "titan_amd.contour", "titan_amd.binding"):
        text = open(importlib.util.find_spec(mod).origin).read()
        assert "from oracle" not in text and "import oracle" not in text
