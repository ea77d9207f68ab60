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
