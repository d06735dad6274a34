"""The C-ABI library loads and exports every entry point include/gpsat.h declares
(no compute calls here: this file runs without a GPU)."""
import ctypes
import os
import re

import pytest

import gpupsat_amd as G


def declared_symbols():
    text = open(G.HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gpsat_[a-z0-9_]+)\s*\(", text)))


def test_library_built_and_loads():
    assert os.path.exists(G.LIB_PATH)
    G.lib()


def test_every_declared_symbol_is_exported():
    syms = declared_symbols()
    assert len(syms) > 10
    lib = G.lib()
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_version_and_errors():
    assert G.lib().gpsat_version() == b"v0.0.1"
    # an argument error sets the message and returns GPSAT_ERR_ARG without a device
    rc = G.lib().gpsat_batch_info(None, None, None, None, None)
    assert rc == -1
    assert b"null" in G.lib().gpsat_last_error()


def test_no_device_is_an_error_not_a_fallback():
    if G.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(G.GpsatError):
        G.GpuBatch(G.Batch.from_formulas([G.Formula.from_clauses(2, [[1, 3]])]))


def test_struct_mirrors_match_the_header(tmp_path):
    """The ctypes mirrors of gpsat_solve_opts / gpsat_solve_info have the C layout (size and
    every field offset), checked against the header compiled with the host compiler."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no host C compiler")
    lines = []
    for cname, py in (("gpsat_solve_opts", G.SolveOpts), ("gpsat_solve_info", G.SolveInfo)):
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    src = tmp_path / "layout.c"
    src.write_text("#include <stdio.h>\n#include <stddef.h>\n#include <stdint.h>\n"
                   f'#include "{G.HEADER_PATH}"\nint main(void) {{\n' + "\n".join(lines) + "\nreturn 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-o", str(exe), str(src)], check=True)
    got = {}
    for ln in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        cname, field, val = ln.split()
        got[(cname, field)] = int(val)
    for cname, py in (("gpsat_solve_opts", G.SolveOpts), ("gpsat_solve_info", G.SolveInfo)):
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)


def test_integration_example_compiles(tmp_path):
    """The C++ caller example of the repository's integration guide (its first cpp block) is
    type-correct against include/gpsat.h (compiled with the host compiler inside a function
    that declares its inputs; not run)."""
    import shutil
    import subprocess
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    text = open(os.path.join(os.path.dirname(G.HEADER_PATH), "..", "INTEGRATION.md")).read()
    block = re.search(r"```cpp\n(.*?)```", text, flags=re.S).group(1)
    body = "\n".join(l for l in block.splitlines() if not l.startswith("#include"))
    src = tmp_path / "example.cpp"
    src.write_text('#include <cstdio>\n#include <cstdint>\n#include <vector>\n#include "gpsat.h"\n'
                   "int example(const int32_t* lits, const int64_t* clause_off, int64_t n_clauses, int32_t n_vars)\n"
                   "{\n" + body + "\nreturn status;\n}\n")
    subprocess.run([cxx, "-std=c++17", "-fsyntax-only", "-Wall", "-I", os.path.dirname(G.HEADER_PATH), str(src)],
                   check=True)
