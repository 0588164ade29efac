"""hr_sincsq_reduce_static (no workspace, no allocation) against
hr_sincsq_reduce with a caller workspace: bit for bit, on two calls in a row
(the second smaller, so the static scratch holds the first call's partials
when it starts), and a size beyond the static scratch refused (GPU)."""

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _args(E_r, a, w, n_E, out):
    from pyrite_amd import _lib

    args = _lib.SincsqArgs()
    args.n_lines, args.E_r, args.a, args.w = int(E_r.numel()), E_r.data_ptr(), a.data_ptr(), w.data_ptr()
    args.n_E, args.E_grid, args.uniform, args.E0, args.dE = int(n_E), None, 1, 1000.0, 47.9
    args.out = out.data_ptr()
    return args


def _lines(rng, n, device):
    import torch

    E_r = torch.as_tensor(rng.uniform(1000.0, 50000.0, n), device=device)
    a = torch.as_tensor(rng.uniform(1e-3, 5e-2, n), device=device)
    w = torch.as_tensor(rng.uniform(0.0, 1.0, n).astype(np.float32), device=device)
    w[::7] = 0.0  # zero-weight lines are skipped by both
    return E_r, a, w


def test_static_scratch_equals_workspace_route(device):
    import torch

    from pyrite_amd import _lib

    lib = _lib.load()
    rng = np.random.default_rng(11)
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    for n, n_E in ((300_000, 1024), (40_000, 700)):  # second call smaller: stale partials in the array
        print(f"no-alloc sinc^2: {n} lines x {n_E} bins", flush=True)
        E_r, a, w = _lines(rng, n, device)
        want = torch.zeros(n_E, dtype=torch.float64, device=device)
        got = torch.full((n_E,), 0.25, dtype=torch.float64, device=device)  # both accumulate into out
        want += 0.25
        ref = _args(E_r, a, w, n_E, want)
        ref.ws = _lib.DeviceScratch(device).get(_lib.workspace_bytes("hr_sincsq_workspace_bytes", C.byref(ref)))
        _lib.check(lib.hr_sincsq_reduce(C.byref(ref), stream), "hr_sincsq_reduce")
        _lib.check(lib.hr_sincsq_reduce_static(C.byref(_args(E_r, a, w, n_E, got)), stream),
                   "hr_sincsq_reduce_static")
        torch.cuda.synchronize()
        assert float(want.sum()) > 0.25 * n_E
        np.testing.assert_array_equal(got.cpu().numpy(), want.cpu().numpy())


def test_static_scratch_refuses_sizes_beyond_its_array(device):
    import torch

    from pyrite_amd import _lib

    lib = _lib.load()
    n_E = 1 << 22  # partials of 4M bins exceed the 32 MiB array; refused before any launch
    E_r, a, w = _lines(np.random.default_rng(1), 1000, device)
    out = torch.zeros(n_E, dtype=torch.float64, device=device)
    assert lib.hr_sincsq_reduce_static(C.byref(_args(E_r, a, w, n_E, out)), None) == _lib.HR_ERR_INVALID
    assert "static" in lib.hr_last_error().decode()
    torch.cuda.synchronize()
    assert float(out.abs().sum()) == 0.0
