"""Builds libhiprite.so (gfx950) in-tree with hipcc.

``python -m pyrite_amd.build`` compiles every ``csrc/*.hip`` translation unit
for ``--offload-arch=gfx950`` and links ``pyrite_amd/_lib/libhiprite.so``.
Packed-FP32 instruction selection is disabled on every unit
(``-target-feature -packed-fp32-ops``).  Units whose arithmetic is compared
against the FP64 oracle step by step are compiled with ``-ffp-contract=off``
(the reproducible math mode); the others keep hipcc's default contraction.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "_lib"
LIB = LIBDIR / "libhiprite.so"
BUILD = PKG / "_build"

ARCH = os.environ.get("HIPRITE_ARCH", "gfx950")
# Units checked step by step against the FP64 oracle: no fused multiply-add.
REPRODUCIBLE = {"transport.hip", "vr.hip", "cone.hip"}


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; set HIPCC or install ROCm")


def _compile(src: Path, build_dir: Path = BUILD, extra=()) -> Path:
    obj = build_dir / (src.stem + ".o")
    flags = [
        _hipcc(),
        "-c",
        "-fPIC",
        "-O3",
        "-std=c++17",
        f"--offload-arch={ARCH}",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", str(ROOT / "include"),
        "-I", str(CSRC),
        "-Wall",
        "-Wno-unused-function",
    ]
    if src.name in REPRODUCIBLE:
        flags.append("-ffp-contract=off")
    flags += list(extra)
    flags += [str(src), "-o", str(obj)]
    res = subprocess.run(flags, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = True, out: Path | None = None, extra_flags=()) -> Path:
    """Compile every unit and link the shared library; returns its path.

    ``out`` and ``extra_flags`` build a variant elsewhere (e.g. with
    ``-DHR_TRANSPORT_WAVES_PER_EU=3``) without touching the in-tree library."""
    lib = LIB if out is None else Path(out)
    build_dir = BUILD if out is None else lib.parent / "_build"
    build_dir.mkdir(parents=True, exist_ok=True)
    lib.parent.mkdir(parents=True, exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    with ThreadPoolExecutor(max_workers=min(8, len(sources))) as pool:
        objs = list(pool.map(lambda src: _compile(src, build_dir, extra_flags), sources))
    LIB_ = lib
    tmp = LIB_.with_suffix(".so.tmp")
    cmd = [_hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs), "-L/opt/rocm/lib",
           "-lrocprofiler-sdk-roctx", "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB_)
    if verbose:
        print(f"built {LIB_}")
    return LIB_


if __name__ == "__main__":
    build()
    sys.exit(0)
