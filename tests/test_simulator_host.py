"""Host-side validation of the simulator callers (raised before any GPU call),
restating invariantq/simulator/statevector.py:31-44 and density_matrix.py:26-31
and the reference tests tests/test_core_diagnostics.py:34-41."""

import pytest

from invariantq_amd.simulator import DensityMatrixSimulator, StateVectorSimulator, checked_power_of_two


@pytest.mark.parametrize("shots", [0, -3, True, 2.5])
def test_shots_must_be_positive_integer(shots):
    with pytest.raises(ValueError, match="positive integer"):
        StateVectorSimulator().run(1, [("gate", "h", [0], None)], shots=shots)


def test_seed_and_empty_circuit():
    with pytest.raises(ValueError, match="non-negative"):
        StateVectorSimulator().run(1, [], shots=4, seed=-1)
    with pytest.raises(ValueError, match="0 qubits"):
        StateVectorSimulator().run(0, [], shots=4)


def test_memory_limits():
    with pytest.raises(MemoryError, match="above configured limit"):
        StateVectorSimulator(max_memory_bytes=16).run(1, [("gate", "h", [0], None)], shots=2)
    with pytest.raises(MemoryError, match="n > 14"):
        DensityMatrixSimulator().run(15, [], shots=2)
    assert checked_power_of_two(3, 16, copies=2) == 256
    with pytest.raises(MemoryError):
        checked_power_of_two(62, 16)
