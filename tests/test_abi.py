"""libhiprite.so loads and exports every symbol include/hiprite.h declares.

No compute call: this runs on machines without a GPU.
"""

import ctypes
import re
from pathlib import Path

import pytest

from pyrite_amd import _lib

HEADER = Path(__file__).resolve().parent.parent / "include" / "hiprite.h"


def declared_symbols():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hr_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "hr_transport_run" in syms and "hr_stream_keys" in syms


def test_library_exports_every_declared_symbol():
    if not _lib.LIB_PATH.exists():
        pytest.fail("libhiprite.so is not built; run python -m pyrite_amd.build")
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, f"not exported: {missing}"
    # the ctypes table covers the same set
    assert sorted(_lib.SIGNATURES) == declared_symbols()


def test_version_string_without_gpu():
    lib = _lib.load()
    assert lib.hr_version().decode().startswith("hiprite")
    assert lib.hr_device_count() >= 0


def test_product_does_not_import_oracle():
    root = Path(__file__).resolve().parent.parent / "pyrite_amd"
    for path in root.rglob("*.py"):
        src = path.read_text()
        assert "oracle" not in re.findall(r"^\s*(?:from|import)\s+(\w+)", src, flags=re.M), path
