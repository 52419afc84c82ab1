"""The simulator callers on the GPU engine, against the reference's own
simulator-level tests (tests/test_statevector.py:14-60,
tests/test_core_math.py:95-115, tests/test_independent_reference.py:44-61) and
the oracle."""

import numpy as np
import pytest

from invariantq_amd import gates, workloads
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    from invariantq_amd import simulator
    return simulator


GHZ = [("gate", "h", [0], None), ("gate", "cx", [0, 1], None), ("gate", "cx", [1, 2], None)]


def test_statevector_ghz(S):
    res = S.StateVectorSimulator().run(3, GHZ, shots=2000, seed=42)
    assert set(res.counts) == {"000", "111"}
    assert sum(res.counts.values()) == 2000
    assert abs(res.probabilities["000"] - 0.5) < 1e-12 and abs(res.probabilities["111"] - 0.5) < 1e-12
    assert res.metadata["core"]["implementation"] == "hip-gfx950"
    assert abs(res.entanglement_entropy - 1.0) < 1e-10  # half cut of GHZ-3 = one bit


def test_counts_match_oracle_samples(S, oracle_lib):
    n = 6
    ops = workloads.random_u3_cx(n, 4, seed=5)
    res = S.StateVectorSimulator().run(n, ops, shots=500, seed=1234)
    ref = oracle_lib.statevector(n)
    for op in ops:
        (ref.apply_1q(op.qubits[0], op.matrix) if len(op.qubits) == 1 else ref.apply_2q(op.qubits[0], op.qubits[1], op.matrix))
    expect = {}
    for s in ref.sample(500, 1234):
        k = f"{int(s):0{n}b}"
        expect[k] = expect.get(k, 0) + 1
    assert res.counts == expect
    assert np.max(np.abs(res.statevector - ref.get_amplitudes())) < 1e-12


def test_non_unitary_isolation(S):
    broken = [("gate", "broken", [0], np.array([[2.0, 0.0], [0.0, 1.0]]))]
    with pytest.raises(ValueError, match="finite unitary"):
        S.StateVectorSimulator().run(1, broken, shots=4)
    res = S.NoisyStatevectorSimulator().run(1, broken, shots=4, seed=1)
    assert res.metadata["allow_non_unitary"] is True
    assert res.metadata["state_norm"] == pytest.approx(4.0)


def test_density_matrix_bell_and_noise(S):
    bell = [("gate", "h", [0], None), ("gate", "cx", [0, 1], None)]
    res = S.DensityMatrixSimulator().run(2, bell, shots=1000)
    assert abs(res.metadata["trace"] - 1) < 1e-9 and abs(res.metadata["purity"] - 1) < 1e-9
    noisy = [("gate", "h", [0], None), ("noise", "depolarizing", [0], [0.2])]
    res = S.DensityMatrixSimulator().run(1, noisy, shots=1000)
    assert abs(res.metadata["trace"] - 1) < 1e-9 and res.metadata["purity"] < 1.0
    # rho = |psi><psi| (tests/test_independent_reference.py:44-61)
    circ = [("gate", "h", [0], None), ("gate", "rx", [1], gates.rx(0.4)), ("gate", "cx", [0, 1], None),
            ("gate", "ry", [2], gates.ry(0.6)), ("gate", "cx", [1, 2], None)]
    sv = S.StateVectorSimulator().run(3, circ, shots=10)
    dm = S.DensityMatrixSimulator().run(3, circ, shots=10)
    assert np.linalg.norm(np.outer(sv.statevector, sv.statevector.conj()) - dm.density_matrix) < 1e-10


def test_noise_model_and_two_qubit_kraus(S):
    # 2q depolarizing-like channel as 16 Kraus operators on (0, 1) after every CX
    rng = np.random.default_rng(3)
    ks1 = gates.depolarizing(0.1)
    ks2 = [np.kron(a, b) for a in ks1 for b in ks1]  # product channel: Kraus set of the pair
    model = {"cx": ks2}
    circ = [("gate", "h", [0], None), ("gate", "cx", [0, 1], None)]
    res = S.DensityMatrixSimulator(noise_model=model).run(2, circ, shots=100)
    # reference: apply the 1q channel on each qubit of the Bell state
    psi = np.array([1, 0, 0, 1]) / np.sqrt(2)
    rho = np.outer(psi, psi.conj())
    full = lambda K, q: orc.full_operator(2, [q], K)
    for q in (0, 1):
        rho = sum(full(K, q) @ rho @ full(K, q).conj().T for K in ks1)
    assert np.max(np.abs(res.density_matrix - rho)) < 1e-12
    assert abs(res.metadata["trace"] - 1) < 1e-12
