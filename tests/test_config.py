"""Run-config mirror: the reference's field names, validation and lowering (CPU)."""

import pytest

from pyrite_amd.config import ExecutionConfig, TransportKernelTuning, lower


def test_defaults_lower_to_the_batched_route():
    from pyrite_amd.transport.core import DEFAULT_CONFIG

    lo = lower(ExecutionConfig())
    assert lo.schedule == "full-history" and lo.precision == "fp64"
    assert lo.transport_config == DEFAULT_CONFIG


def test_tuning_fields_lower_onto_the_batching_policy():
    t = TransportKernelTuning(schedule="continuation", continuation_rows=48, electron_batch=4096, scratch_mib=512,
                              kernel_threads=32, precision="mixed")
    lo = lower(ExecutionConfig(transport=t))
    assert lo.schedule == "continuation" and lo.rows_per_round == 48 and lo.precision == "mixed"
    assert lo.transport_config.scratch_budget_bytes == 512 << 20
    assert lo.transport_config.max_batch == 4096
    assert lo.transport_config.kernel_threads == 64  # 32 lanes run as one 64-lane wave
    assert lower(ExecutionConfig(stream_block_electrons=10_000)).schedule == "continuation"


@pytest.mark.parametrize("bad", [
    dict(schedule="sometimes"), dict(precision="fp16"), dict(kernel_threads=48), dict(scratch_mib=0),
    dict(electron_batch=True), dict(continuation_rows=16),
])
def test_tuning_refuses_bad_fields(bad):
    with pytest.raises(ValueError):
        TransportKernelTuning(**bad)


def test_execution_config_refuses_bad_or_foreign_fields():
    with pytest.raises(ValueError):
        ExecutionConfig(transport_core="lockstep")
    with pytest.raises(ValueError):
        ExecutionConfig(stream_block_electrons=10, stream_memory_mib=10)
    with pytest.raises(ValueError):
        ExecutionConfig(math="turbo")
    with pytest.raises(TypeError):
        ExecutionConfig(transport={"schedule": "auto"})
    with pytest.raises(NotImplementedError):
        ExecutionConfig(stream_spill_mib=0)
