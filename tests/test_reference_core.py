"""The reference-positional core adapter maps every tuple member onto the same
``hr_transport_args`` field the package's own driver sets (CPU: no launch)."""

import numpy as np
import pytest

from tests.transport_cases import MULTILAYER


def _setup():
    import torch

    from pyrite_amd.transport.core import DeviceTables, TransportControl, alloc_rows

    lt, lut = MULTILAYER.tables()
    dev = torch.device("cpu")
    tabs = DeviceTables(lt, lut, dev)
    ctl = TransportControl(max_steps=777, energy_model=1, max_dE_frac=0.05, straggle_on=True)
    n, cap = 50, 16
    keys = torch.arange(n, dtype=torch.int64).to(torch.uint64)
    state = (torch.ones(n, dtype=torch.uint8), torch.zeros(n, dtype=torch.float64),
             torch.zeros((n, 3), dtype=torch.float64), torch.zeros((n, 3), dtype=torch.float64),
             torch.full((n,), 40.0, dtype=torch.float64), torch.full((n,), 5.0, dtype=torch.float64))
    rows = alloc_rows(n - 10, cap, True, dev)
    out = (torch.zeros(n - 10, dtype=torch.int64), torch.zeros(n - 10, dtype=torch.int8))
    stragg = torch.zeros(n, dtype=torch.float64)
    return tabs, ctl, (10, n - 10, cap, keys), state, rows, out, stragg


def test_positional_call_maps_onto_the_native_arguments():
    from pyrite_amd import _lib
    from pyrite_amd.transport.core import transport_args
    from pyrite_amd.transport.reference_core import build_args, reference_arguments

    tabs, ctl, run, state, rows, out, stragg = _setup()
    pos_args = reference_arguments(tabs, ctl, run, state, rows, out, stragg)
    a = build_args(*pos_args)
    b = transport_args(tabs, ctl, run, state, rows, out, (ctl.straggle_on, stragg))
    for name, _ in _lib.TransportArgs._fields_:
        if name in ("work", "hard", "cone", "elsepa") or name.endswith("_f"):
            continue  # claim counter and FP32 copies are the caller's; modes off in both
        assert getattr(a, name) == getattr(b, name), name
    assert a.max_steps == 777 and a.max_dE_frac == 0.05 and a.e_start == 10 and a.cap == 16


def test_positional_call_fails_closed():
    from pyrite_amd.transport.reference_core import build_args, make_reference_core, reference_arguments

    tabs, ctl, run, state, rows, out, stragg = _setup()
    run_, control, geometry, lut_args, materials, st, segments, pe_out, straggling = reference_arguments(
        tabs, ctl, run, state, rows, out, stragg)
    with pytest.raises(NotImplementedError, match="ELSEPA"):
        build_args(run_, (777, 0, 2, 1, 0.0), geometry, lut_args, materials, st, segments, pe_out, straggling)
    with pytest.raises(NotImplementedError, match="SBETHE"):
        build_args(run_, control, geometry, lut_args, materials[:5] + (1,), st, segments, pe_out, straggling)
    with pytest.raises(TypeError, match="seg_len must be float64"):
        bad = list(segments)
        bad[2] = bad[2].float()
        build_args(run_, control, geometry, lut_args, materials, st, tuple(bad), pe_out, straggling)
    with pytest.raises(ValueError, match="12 midpoint columns"):
        build_args(run_, control, geometry, lut_args, materials, st, segments[:7], pe_out, straggling)
    with pytest.raises(NotImplementedError, match="inelastic_args"):
        make_reference_core()(run_, control, geometry, lut_args, materials, st, segments, pe_out, straggling,
                              (1, 2, 3))
    np.testing.assert_equal(len(segments), 12)


def test_mott_code_runs_the_lut_it_is_given():
    """The reference's LUT kernel takes ``elastic_model_code`` without branching
    on it (the LUT carries the model), so code 1 maps onto the same native
    arguments as code 0."""
    from pyrite_amd import _lib
    from pyrite_amd.transport.reference_core import build_args, reference_arguments

    tabs, ctl, run, state, rows, out, stragg = _setup()
    run_, control, geometry, lut_args, materials, st, segments, pe_out, straggling = reference_arguments(
        tabs, ctl, run, state, rows, out, stragg)
    a = build_args(run_, control, geometry, lut_args, materials, st, segments, pe_out, straggling)
    mott = (control[0], control[1], 1, control[3], control[4])
    b = build_args(run_, mott, geometry, lut_args, materials, st, segments, pe_out, straggling)
    for name, _ in _lib.TransportArgs._fields_:
        if name in ("work", "hard", "cone", "elsepa"):
            continue
        assert getattr(a, name) == getattr(b, name), name
