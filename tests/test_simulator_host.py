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


def test_batched_unitarity_check_matches_per_gate_allclose():
    """The gates' unitarity test made for all gates of one size at once gives
    the per-gate decision of statevector.py:55-65 (finite, and
    np.allclose(U^H U, I, atol=1e-10, rtol=1e-10)), near the tolerance too."""
    import numpy as np
    from invariantq_amd import workloads
    from invariantq_amd.simulator import _unitary_flags
    rng = np.random.default_rng(1)
    cases = []
    for k in range(600):
        d = (2, 4, 8)[k % 3]
        U = workloads.haar_unitary(rng, d)
        U = U + 10.0 ** rng.uniform(-13, -8) * (rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d)))
        if k % 97 == 0:
            U[0, 0] = np.nan
        cases.append(("gate", "u", list(range(d.bit_length() - 1)), U))
    cases.insert(5, ("reset", None, [0], None))
    cases.insert(9, ("gate", "bad", [0], np.eye(4)))  # wrong size for one qubit
    got = _unitary_flags(cases)
    want = [True if k != "gate" else bool(np.asarray(m).shape == (1 << len(q), 1 << len(q))
                                          and np.isfinite(m).all()
                                          and np.allclose(np.conj(m).T @ m, np.eye(m.shape[0]), atol=1e-10, rtol=1e-10))
            for k, _, q, m in cases]
    assert got == want
    assert 0 < sum(want) < len(want)
