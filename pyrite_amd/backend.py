"""The ``hip`` backend: selection, device identity and OOM recognition.

Mirrors the reference selector (``src/pyrite/_backend.py:30,358-464``): the
valid set gains ``"hip"``; an explicit request never falls back, only
``"auto"`` may.  This package implements only the hip device path, so a
request for another backend is an error here (the reference serves those).
Device memory and streams are torch.cuda (ROCm); the kernels are
libhiprite.so's.
"""

from __future__ import annotations

import os
from dataclasses import dataclass

from . import _lib

VALID_BACKENDS = ("auto", "cpu", "cuda", "rocm", "sycl", "hip")
ENV = "PYRITE_MC_BACKEND"


class BackendError(RuntimeError):
    """Base error for backend selection (``_backend.py:33``)."""


class BackendUnavailableError(BackendError):
    """Requested accelerator backend is unavailable (``_backend.py:37``)."""


@dataclass(frozen=True)
class DeviceInfo:
    """Device identity used by run planning (``_backend.py:45-53``)."""

    backend: str
    vendor: str
    name: str
    total_memory_bytes: int | None
    supports_fp64: bool
    arch: str


@dataclass
class HipBackend:
    name: str
    device: DeviceInfo
    torch_device: object

    def is_oom_error(self, error: BaseException) -> bool:
        """OOM from torch's allocator or from libhiprite (``_backend.py:179-205``)."""
        return is_oom_error(error)

    def synchronize(self) -> None:
        import torch

        torch.cuda.synchronize(self.torch_device)


def is_oom_error(error: BaseException) -> bool:
    """Whether ``error`` is a device out-of-memory failure (``_backend.py:179-205``):
    torch's allocator error, libhiprite's ``HR_ERR_OOM``, or a message naming it."""
    if isinstance(error, (_lib.HipriteOOM, MemoryError)):
        return True
    try:
        import torch

        if isinstance(error, torch.cuda.OutOfMemoryError):
            return True
    except ImportError:  # pragma: no cover
        pass
    msg = str(error).lower()
    return "out of memory" in msg or "hiperroroutofmemory" in msg.replace("_", "")


def _probe_hip() -> HipBackend:
    try:
        import torch
    except ImportError as exc:  # pragma: no cover
        raise BackendUnavailableError(f"torch is not importable: {exc}") from exc
    if not getattr(torch.version, "hip", None):
        raise BackendUnavailableError("torch is not a ROCm build")
    if not torch.cuda.is_available():
        raise BackendUnavailableError("no HIP device is visible")
    try:
        _lib.load()
    except _lib.HipriteUnavailable as exc:
        raise BackendUnavailableError(str(exc)) from exc
    dev = torch.device("cuda", torch.cuda.current_device())
    props = torch.cuda.get_device_properties(dev)
    arch = getattr(props, "gcnArchName", "") or ""
    info = DeviceInfo("hip", "amd", props.name, int(props.total_memory), True, arch)
    return HipBackend("hip", info, dev)


def select_backend(requested: str | None = None) -> HipBackend:
    """Resolve ``requested`` (default ``$PYRITE_MC_BACKEND`` or ``"auto"``).

    ``"hip"`` (and ``"rocm"``, which names the same device under ROCm PyTorch)
    must resolve to the device or raise; ``"auto"`` resolves to the device
    when present and otherwise raises too, since this package has no host
    path of its own.
    """
    if requested is None:
        requested = os.environ.get(ENV, "auto")
    requested = requested.strip().lower()
    if requested not in VALID_BACKENDS:
        raise BackendUnavailableError(
            f"{ENV} must be one of {', '.join(VALID_BACKENDS)}; got {requested!r}"
        )
    if requested in ("cpu", "cuda", "sycl"):
        raise BackendUnavailableError(
            f"backend {requested!r} is served by the reference PyRITE, not by pyrite_amd"
        )
    return _probe_hip()


_BACKEND: HipBackend | None = None


def require_device() -> HipBackend:
    """The resolved hip backend (cached); raises when no device or library."""
    global _BACKEND
    if _BACKEND is None:
        _BACKEND = select_backend()
    return _BACKEND
