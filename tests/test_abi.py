"""The C-ABI library loads and exports every function include/elfel_hip.h declares
(no device calls: runs without a GPU)."""
import ctypes
import os
import re

import elfel_hip as eh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    with open(os.path.join(ROOT, "include", "elfel_hip.h")) as fh:
        text = fh.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(elfel_hip_\w+)\s*\(", text)))


def test_header_declares_the_python_binding_set():
    assert declared_functions() == sorted(eh._lib.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(eh._lib.LIBPATH)
    missing = [f for f in declared_functions() if not hasattr(lib, f)]
    assert not missing, f"symbols missing from libelfel_hip.so: {missing}"


def test_library_is_built_for_gfx950():
    with open(eh._lib.LIBPATH, "rb") as fh:
        blob = fh.read()
    assert b"gfx950" in blob
