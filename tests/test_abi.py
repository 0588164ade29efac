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


STRUCTS = {
    "hr_transport_args": "TransportArgs",
    "hr_line_args": "LineArgs",
    "hr_sincsq_args": "SincsqArgs",
    "hr_brem_args": "BremArgs",
    "hr_mlp": "MlpNet",
    "hr_vr_window_args": "VrWindowArgs",
    "hr_vr_spawn_args": "VrSpawnArgs",
    "hr_vr_gather_args": "VrGatherArgs",
    "hr_beam_args": "BeamArgs",
    "hr_window_args": "WindowArgs",
    "hr_coherent_args": "CoherentArgs",
    "hr_coherent_field_args": "CoherentFieldArgs",
    "hr_coherent_grouped_args": "CoherentGroupedArgs",
    "hr_queue_config": "QueueConfig",
    "hr_queue_launch": "QueueLaunch",
    "hr_continuation": "Continuation",
    "hr_dose_args": "DoseArgs",
    "hr_workspace": "Workspace",
    "hr_hard_events": "HardEvents",
    "hr_cone_split": "ConeSplitArgs",
    "hr_cone_collect_args": "ConeCollectArgs",
    "hr_elsepa": "ElsepaArgs",
}


def test_ctypes_layouts_match_the_c_header(tmp_path):
    """Every field of every ctypes mirror sits at the C compiler's offset."""
    import shutil
    import subprocess

    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "hiprite.h"', "int main(void) {"]
    for cname, pyname in STRUCTS.items():
        cls = getattr(_lib, pyname)
        lines.append(f'printf("{pyname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{pyname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", f"-I{HEADER.parent}", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    for row in filter(None, out):
        pyname, field, value = row.split()
        cls = getattr(_lib, pyname)
        got = ctypes.sizeof(cls) if field == "size" else getattr(cls, field).offset
        assert got == int(value), (pyname, field, got, value)
