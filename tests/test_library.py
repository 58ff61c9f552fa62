"""The C-ABI library loads and exports every symbol include/ptengine.h
declares (no compute calls, CPU only)."""

import os
import re

from oqupy_amd.backends import engine

HEADER = os.path.join(os.path.dirname(os.path.dirname(__file__)), "include",
                      "ptengine.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pte_\w+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_functions() == engine.exported_symbols()


def test_library_exports_every_declared_symbol():
    lib = engine.load_library()
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert lib.pte_version().startswith(b"ptengine")


def test_backend_config_rejects_other_backends():
    import pytest
    with pytest.raises(ValueError):
        engine.parse_backend_config({"backend": "tensornetwork"})
    assert engine.parse_backend_config(None)["backend"] == "mi355x"
