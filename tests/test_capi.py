"""CPU checks of the C ABI: the library loads, exports every symbol that
include/txa.h declares, and reports errors (no compute calls without a GPU)."""
import ctypes
import os

import pytest

import tianxin_amd as txa
from tianxin_amd import build as txbuild


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(txa.LIB_PATH):
        txbuild.build()
    return ctypes.CDLL(txa.LIB_PATH)


def test_header_declares_the_boundary():
    names = txa.header_symbols()
    for n in ("txa_create", "txa_destroy", "txa_last_error", "txa_mesh_brick", "txa_dofmap_build",
              "txa_dofmap_copy", "txa_graph_build", "txa_graph_copy", "txa_graph_device", "txa_assemble"):
        assert n in names


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in txa.header_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_version_string(lib):
    lib.txa_version.restype = ctypes.c_char_p
    assert b"gfx950" in lib.txa_version()


def test_null_context_is_rejected(lib):
    lib.txa_sync.argtypes = [ctypes.c_void_p]
    assert lib.txa_sync(None) == txa.TXA_ERR_INVALID
    lib.txa_assemble.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    assert lib.txa_assemble(None, 1, None, None) == txa.TXA_ERR_INVALID


def test_bench_refuses_more_than_one_gpu():
    import json
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]


def _unbound_names(path):
    """Names a Python file reads that nothing in it binds (and that are not builtins)."""
    import ast
    import builtins
    tree = ast.parse(open(path).read())
    bound, loaded = set(dir(builtins)) | {"__file__", "__name__"}, {}
    for node in ast.walk(tree):
        if isinstance(node, ast.Name):
            if isinstance(node.ctx, ast.Load):
                loaded.setdefault(node.id, node.lineno)
            else:
                bound.add(node.id)
        elif isinstance(node, (ast.FunctionDef, ast.AsyncFunctionDef, ast.ClassDef)):
            bound.add(node.name)
        elif isinstance(node, ast.arg):
            bound.add(node.arg)
        elif isinstance(node, (ast.Import, ast.ImportFrom)):
            bound.update((a.asname or a.name).split(".")[0] for a in node.names)
        elif isinstance(node, ast.ExceptHandler) and node.name:
            bound.add(node.name)
    return {n: ln for n, ln in loaded.items() if n not in bound}


def test_bench_reads_no_undefined_name():
    # every code path of bench.py (the GPU ones included) only reads names it defines or imports
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    for f in ("bench.py", "__graft_entry__.py"):
        assert _unbound_names(os.path.join(root, f)) == {}, f


def test_ctypes_mirrors_match_the_header_layout(tmp_path):
    # txa_assembly_args and txa_serve_request: every field offset and the size, C compiler vs ctypes
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    structs = {"txa_assembly_args": txa.AssemblyArgs, "txa_serve_request": txa.ServeRequest}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "txa.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'  printf("%zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'  printf("%zu\\n", offsetof({cname}, {fname}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(root, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = []
    for py in structs.values():
        want.append(ctypes.sizeof(py))
        want += [getattr(py, fname).offset for fname, _ in py._fields_]
    assert got == want
