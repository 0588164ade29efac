"""Depth-dose tally of device transport rows (``hr_depth_dose``).

The energy each row loses is spread over the depth bins its flight crosses,
and cut-off histories leave their residual energy at their last row's end, so
the profile and the escaping electrons' energy add up to Ne E0 exactly.
"""

from __future__ import annotations

import ctypes as C

from .. import _lib


def depth_dose(out, z0, dz, rows, n, slots=1, seg_count=None, weight=None, scale=1.0, stream=None):
    """Accumulate into ``out`` ([n_bins] FP64 device tensor) the dose of ``n``
    particles' rows held slot-major (``slots`` per particle, ``seg_count`` rows
    each; ``slots=1`` and ``seg_count=None`` for compacted rows)."""
    import torch

    a = _lib.DoseArgs()
    a.n, a.rows = int(n), int(slots)
    a.seg_count = 0 if seg_count is None else seg_count.data_ptr()
    a.seg_mid, a.seg_dir, a.seg_len = rows["seg_mid"].data_ptr(), rows["seg_dir"].data_ptr(), rows["seg_len"].data_ptr()
    a.seg_E, a.seg_E_end = rows["seg_E"].data_ptr(), rows["seg_E_end"].data_ptr()
    a.seg_event = rows["seg_event"].data_ptr() if rows.get("seg_event") is not None else 0
    a.weight = 0 if weight is None else weight.data_ptr()
    a.scale, a.z0, a.dz = float(scale), float(z0), float(dz)
    a.n_bins = int(out.shape[0])
    a.out = out.data_ptr()
    if stream is None:
        stream = torch.cuda.current_stream().cuda_stream
    _lib.check(_lib.load().hr_depth_dose(C.byref(a), C.c_void_p(stream)), "hr_depth_dose")
    return out
