"""The C-ABI library loads and exports every entry point its header
declares; no compute calls (CPU only)."""
import ctypes
import os
import re

import pytest

from seispie_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header="seispie_hip.h"):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(sp_[a-z_]+)\s*\(", txt)))


def test_library_is_built():
    assert os.path.exists(_lib.LIB_PATH), "run __graft_entry__.build() first"


def test_every_declared_symbol_is_exported_and_bound():
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.exported_symbols())
    _lib.load()


def test_network_symbols_are_exported_and_bound():
    from seispie_amd import cnn
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared("seispie_cnn.h")
    assert len(names) >= 10
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(cnn.exported_symbols())
    cnn._load()


def test_library_is_built_for_gfx950_only():
    blob = open(_lib.LIB_PATH, "rb").read()
    targets = set(re.findall(rb"amdgcn-amd-amdhsa--(gfx[0-9a-f]+)", blob))
    assert targets == {b"gfx950"}, targets


def test_context_without_a_device_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(_lib.SeispieError):
        _lib.Context(0)
    from seispie_amd import cnn
    with pytest.raises(_lib.SeispieError):
        cnn.Net([2, 4, 1])
