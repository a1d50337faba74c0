"""The C-ABI library loads and exports every symbol include/cfmi.h declares (CPU)."""

import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "cfmi.h")).read()
    return sorted(set(re.findall(r"\b(cfmi_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_header_symbols():
    from cosmoforge_amd import _lib

    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in header_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_header():
    from cosmoforge_amd import _lib

    assert sorted(_lib.SIGNATURES) == header_symbols()


def test_library_version_and_null_context():
    from cosmoforge_amd import _lib

    lib = _lib.load()
    assert lib.cfmi_version() >= 10000
    assert lib.cfmi_destroy(None) == 0
    assert lib.cfmi_synchronize(None) < 0
