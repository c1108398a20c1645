"""The C-ABI library loads and exports every symbol include/phf_hip.h declares
(no GPU needed: no compute calls)."""
import os
import re

from phf_pcc_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"\b(phf_[a-z0-9_]+)\s*\(", text))
    return names


def test_header_declares_entry_points():
    assert {"phf_buildhsg", "phf_localg", "phf_set_hamiltonian"} <= _declared()


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in sorted(_declared()):
        assert hasattr(lib, name), name
    assert set(_lib.EXPORTS) == _declared()


def test_last_error_without_context_is_callable():
    assert isinstance(_lib.load().phf_last_error(None), bytes)
