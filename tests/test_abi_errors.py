"""Error behaviour of the C ABI without a GPU (CPU).

Every entry point validates its arguments before any HIP call and reports a
bad argument as HR_ERR_INVALID with a message in hr_last_error, the way the
reference's host layer fails closed with ValueError before launching
(batching.py, _line_roulette.py:39-52).  Empty work returns HR_OK without
touching the device.  Only argument checks run here: no pointer is ever
dereferenced and no kernel is launched.
"""

import ctypes as C

import pytest

from pyrite_amd import _lib

HR_OK, HR_ERR_INVALID = 0, 1


@pytest.fixture(scope="module")
def lib():
    if not _lib.LIB_PATH.exists():
        pytest.skip("libhiprite.so is not built")
    return _lib.load()


def _err(lib):
    return (lib.hr_last_error() or b"").decode()


def test_null_and_negative_arguments_are_invalid(lib):
    assert lib.hr_transport_run(None, None) == HR_ERR_INVALID
    assert "null args" in _err(lib)
    assert lib.hr_stream_keys(1, 0, -5, None, None) == HR_ERR_INVALID
    assert lib.hr_roulette_uniforms(7, 0, -1, None, None) == HR_ERR_INVALID
    assert lib.hr_sincsq_reduce(None, None) == HR_ERR_INVALID
    assert lib.hr_window_reduce(None, None, None, -1, 16, 0.0, 1.0, 100.0, None, _lib.Workspace(), None) == HR_ERR_INVALID
    assert lib.hr_vr_gather(None, None) == HR_ERR_INVALID


def test_empty_work_is_ok_without_a_device(lib):
    assert lib.hr_stream_keys(1, 0, 0, None, None) == HR_OK
    assert lib.hr_roulette_uniforms(7, 0, 0, None, None) == HR_OK
    assert lib.hr_line_screen(None, 0, 0, 8, 1.0, 0.5, 7, None, None, None) == HR_OK
    b = _lib.BeamArgs()
    b.n, b.start = 0, 0
    assert lib.hr_beam_entry(C.byref(b), None) == HR_OK


def test_line_screen_refuses_a_floor_outside_the_unit_interval(lib):
    for floor in (0.0, -0.5, 1.5):
        assert lib.hr_line_screen(None, 4, 0, 8, 1.0, floor, 7, None, None, None) == HR_ERR_INVALID
        assert "floor" in _err(lib)


def test_continuation_refuses_bad_launch_shapes(lib):
    a, c = _lib.TransportArgs(), _lib.Continuation()
    a.e_start, a.e_count, a.cap = 3, 10, 4
    assert lib.hr_transport_continue(C.byref(a), C.byref(c), None) == HR_ERR_INVALID
    assert "e_start must be 0" in _err(lib)
    a.e_start, a.cap = 0, 0
    assert lib.hr_transport_continue(C.byref(a), C.byref(c), None) == HR_ERR_INVALID
    assert "rows per launch" in _err(lib)
    a.cap, a.n_layers, a.max_el, a.lut_n_energy, a.energy_model = 4, 1, 1, 16, 1
    assert lib.hr_transport_continue(C.byref(a), C.byref(c), None) == HR_ERR_INVALID
    assert "null state" in _err(lib)
    assert lib.hr_transport_continue(C.byref(a), None, None) == HR_ERR_INVALID


def test_transport_refuses_an_unknown_energy_model_or_precision(lib):
    a = _lib.TransportArgs()
    a.e_start, a.e_count, a.cap = 0, 4, 2
    a.n_layers, a.max_el, a.lut_n_energy = 1, 1, 16
    a.energy_model = 2
    assert lib.hr_transport_run(C.byref(a), None) == HR_ERR_INVALID
    assert "energy_model" in _err(lib)


def test_beam_entry_refuses_an_unknown_mode(lib):
    b = _lib.BeamArgs()
    b.n, b.start, b.mode = 4, 0, 7
    assert lib.hr_beam_entry(C.byref(b), None) == HR_ERR_INVALID
    assert "mode 7" in _err(lib)


def test_python_layer_raises_typed_errors(lib):
    with pytest.raises(_lib.HipriteError) as exc:
        _lib.call("hr_stream_keys", 1, -1, 4, None, None)
    assert "negative" in str(exc.value)


def test_coherent_route_validates_before_launching(lib):
    assert lib.hr_coherent_lines(None, None) == HR_ERR_INVALID
    assert "null args" in _err(lib)
    a, la = _lib.CoherentArgs(), _lib.LineArgs()
    a.lines = C.pointer(la)
    a.n, a.rows = 4, 0
    assert lib.hr_coherent_lines(C.byref(a), None) == HR_ERR_INVALID
    assert "bad sizes" in _err(lib)
    a.rows, la.n_dir, la.n_g, la.n_tab, la.n_mu = 8, 1, 1, 16, 1
    a.first, a.ld = 5, 2
    assert lib.hr_coherent_lines(C.byref(a), None) == HR_ERR_INVALID
    assert "record columns" in _err(lib)
    a.first, a.ld = 0, 32
    assert lib.hr_coherent_lines(C.byref(a), None) == HR_ERR_INVALID
    assert "null row" in _err(lib)
    a.n = 0
    assert lib.hr_coherent_lines(C.byref(a), None) == HR_OK
    f = _lib.CoherentFieldArgs()
    f.n_lines, f.ld, f.n_rows, f.n_E = 10, 4, 1, 8
    assert lib.hr_coherent_field(C.byref(f), None) == HR_ERR_INVALID
    f.n_lines = 0
    assert lib.hr_coherent_field(C.byref(f), None) == HR_OK
    assert lib.hr_coherent_finalize(None, 0, 1, None, None, 8, 1.0, None, None) == HR_ERR_INVALID
    assert lib.hr_coherent_finalize(None, 1, 1, None, None, 0, 1.0, None, None) == HR_OK
