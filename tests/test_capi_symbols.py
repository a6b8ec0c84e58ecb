"""CPU checks of the C ABI: the built library loads and exports every symbol
declared in include/*.h, and the ctypes table covers the same set. No device
calls are made."""
import ctypes
import glob
import os
import re

from tests.conftest import PKG_DIR, ROOT


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(pfw_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(os.path.join(PKG_DIR, "libpfw_mi355x.so"))
    names = declared_symbols()
    assert len(names) >= 10
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_signatures_cover_header(pfw):
    from pfw_amd import _capi
    assert {n for n, _, _ in _capi.SIGNATURES} == declared_symbols()
    lib = _capi.load()
    assert lib is _capi.load()


def test_context_create_fails_loudly_without_device(pfw):
    """Without a visible GPU the product raises; it never falls back to the CPU."""
    import pytest
    from pfw_amd import _capi
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is visible")
    with pytest.raises(_capi.DeviceError):
        _capi.Context(0)
