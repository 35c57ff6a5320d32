"""The C-ABI library loads without a GPU and exports exactly what
include/moljax_hip.h declares; the ctypes prototypes match the header."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "moljax_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mjx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_functions():
    names = declared_functions()
    assert "mjx_spectral_pair" in names and "mjx_gs_imex_strang" in names
    assert len(names) >= 8


def test_library_exports_every_declared_symbol():
    from moljax_amd import _hip
    if not _hip.LIB_PATH.exists():
        pytest.fail("libmoljax_hip.so not built: run __graft_entry__.build()")
    lib = _hip.load_library(require_gpu=False)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_prototypes_cover_header():
    from moljax_amd import _hip
    assert sorted(_hip.PROTOTYPES) == declared_functions()


def test_arg_counts_match_header():
    from moljax_amd import _hip
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    for name, (_, args) in _hip.PROTOTYPES.items():
        m = re.search(r"\b" + name + r"\s*\(([^)]*)\)", text)
        assert m, name
        params = [p for p in m.group(1).split(",") if p.strip() and p.strip() != "void"]
        assert len(params) == len(args), (name, len(params), len(args))
