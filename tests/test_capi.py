"""CPU-side checks of the C-ABI library: it loads and exports every symbol
declared in include/*.h, with ctypes signatures for each (no GPU calls)."""
import glob
import os
import re
import subprocess

import pytest

import risgraph_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*(rg_[a-z0-9_]+)\s*\(", text, re.M):
            names.add(m.group(1))
    return names


def test_header_declares_functions():
    names = declared_functions()
    assert "rg_graph_create" in names and "rg_stream_apply" in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    lib = risgraph_amd.load_library()
    out = subprocess.run(["nm", "-D", "--defined-only", risgraph_amd.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    for name in declared_functions():
        assert name in exported, name
        assert getattr(lib, name) is not None


def test_python_signatures_cover_header():
    assert declared_functions() == set(risgraph_amd.SIGNATURES)


def test_library_is_gfx950_code():
    """The embedded offload bundle targets gfx950 (our own build output)."""
    blob = open(risgraph_amd.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


def test_missing_library_fails_loudly(tmp_path):
    import importlib
    mod = importlib.import_module("risgraph_amd")
    saved = mod._lib
    try:
        mod._lib = None
        with pytest.raises(ImportError):
            mod.load_library(str(tmp_path / "nope.so"))
    finally:
        mod._lib = saved
