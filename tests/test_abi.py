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


class _FakeDevice:
    def __init__(self, index):
        self.index = index

    def __str__(self):
        return "cuda:%d" % self.index


class _FakeCudaTensor:
    """Stands in for a torch CUDA tensor (type module under torch) so the
    binding's checks run without a GPU; its data_ptr is never valid."""
    __module__ = "torch.fake"

    def __init__(self, n, device=0):
        self.n, self.device, self.is_cuda, self.dtype = n, _FakeDevice(device), True, "torch.float32"

    def is_contiguous(self):
        return True

    def numel(self):
        return self.n

    def data_ptr(self):
        return 0xDEAD0000


class _RecordingLib:
    """Records every C call; any call at all fails the test."""

    def __init__(self):
        self.calls = []

    def __getattr__(self, name):
        def f(*a):
            self.calls.append(name)
            return 0
        return f


def _offline_context(physics=_lib.PSV, nx=8, nz=6, nt=5, nrec=3):
    c = object.__new__(_lib.Context)
    c._p, c.L, c.device = ctypes.c_void_p(1), _RecordingLib(), 0
    c.physics, c.nt, c.dt = physics, nt, 1e-3
    c.grids = [(nx, nz)]
    c.shots = [(0, 1, nrec)]
    return c


@pytest.mark.parametrize("call,n", [
    ("set_field", 48), ("get_field", 48), ("kernel", 48), ("records", 2 * 3 * 5), ("set_data", 2 * 3 * 5),
    ("set_adjoint_source", 2 * 3 * 5), ("wavefield", 2 * 48), ("smooth", 48), ("set_model_vps", 48)])
def test_device_tensors_of_the_wrong_size_or_device_raise_before_any_library_call(call, n):
    c = _offline_context()
    invoke = {
        "set_field": lambda t: c.set_field(0, _lib.F_MU, t),
        "get_field": lambda t: c.get_field(0, _lib.F_MU, t),
        "kernel": lambda t: c.kernel(0, _lib.K_MU, t),
        "records": lambda t: c.records(0, t),
        "set_data": lambda t: c.set_data(0, t),
        "set_adjoint_source": lambda t: c.set_adjoint_source(0, t),
        "wavefield": lambda t: c.wavefield(0, _lib.WF_V, t),
        "smooth": lambda t: c.smooth(t, 8, 6, 2),
        "set_model_vps": lambda t: c.set_model_vps(0, t, _FakeCudaTensor(48), _FakeCudaTensor(48)),
    }[call]
    for bad in (_FakeCudaTensor(n - 1), _FakeCudaTensor(n + 1), _FakeCudaTensor(n, device=1)):
        with pytest.raises(ValueError):
            invoke(bad)
    assert c.L.calls == []
    invoke(_FakeCudaTensor(n))          # the right size on the right device reaches the library
    assert len(c.L.calls) == 1
