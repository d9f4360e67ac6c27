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


def test_problem_struct_layout_matches_header(tmp_path):
    """The ctypes mirror of elfel_hip_problem has the C layout (size and the
    offsets of every member), checked with the host C compiler."""
    import subprocess
    fields = [f[0] for f in eh._lib.Problem._fields_]
    body = " ".join(f'printf("%zu ", offsetof(elfel_hip_problem, {f}));' for f in fields)
    src = tmp_path / "layout.c"
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "elfel_hip.h"\n'
                   f'int main(void) {{ {body} printf("%zu\\n", sizeof(elfel_hip_problem)); return 0; }}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    want = [getattr(eh._lib.Problem, f).offset for f in fields] + [ctypes.sizeof(eh._lib.Problem)]
    assert got == want
