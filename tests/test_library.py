"""libnemohip.so loads on the CPU host and exports every declared symbol."""
import os
import re

import pytest

from nemo3d_amd import native

HEADER = os.path.join(os.path.dirname(__file__), "..", "include", "nemohip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nh_[a-z_A-Z0-9]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_functions() == sorted(native.SIGNATURES)


def test_library_exports_all_symbols():
    if not os.path.exists(native.LIB_PATH):
        pytest.skip("library not built")
    lib = native.load()
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_no_cpu_fallback(monkeypatch, tmp_path):
    monkeypatch.setattr(native, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(native, "_LIB", None)
    with pytest.raises(RuntimeError):
        native.load()
