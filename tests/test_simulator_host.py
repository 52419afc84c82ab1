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


def test_counts_and_probabilities_match_the_reference_loops():
    """The vectorised result formatting against the per-shot and per-entry
    loops of statevector.py:104-113: same keys, values and key order."""
    import numpy as np
    from invariantq_amd.simulator import _counts, _probabilities
    rng = np.random.default_rng(5)
    for n, shots in ((1, 7), (5, 1000), (12, 3000)):
        samples = rng.integers(0, 1 << n, size=shots, dtype=np.uint64)
        want = {}
        for s in samples:
            key = f"{int(s):0{n}b}"
            want[key] = want.get(key, 0) + 1
        got = _counts(samples, n)
        assert got == want and list(got) == list(want)
        probs = rng.random(1 << n) * 3e-7
        probs[::7] = 1e-7  # the threshold is strict
        wantp = {f"{i:0{n}b}": float(p) for i, p in enumerate(probs) if p > 1e-7}
        gotp = _probabilities(probs, n)
        assert gotp == wantp and list(gotp) == list(wantp)
    assert _counts(np.array([], dtype=np.uint64), 3) == {}
