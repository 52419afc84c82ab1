"""Density-matrix engine: an asymmetric two-qubit Kraus set in both qubit
orders against dense brute force (reference density_matrix.cpp:165-191 takes
the 2q unitary path per operator), the new channels of
invariantq/noise/channels.py:61-102 through the simulator caller, and the
14-qubit noisy workload (2^28 entries, 4 GiB) op by op and fused."""

import numpy as np
import pytest

from invariantq_amd import fusion, gates, workloads
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def random_rho(rng, d):
    a = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
    rho = a @ a.conj().T
    return rho / np.trace(rho)


def kraus_dense(n, qs, ks, rho):
    out = np.zeros_like(rho)
    for K in ks:
        F = orc.full_operator(n, qs, K)
        out += F @ rho @ F.conj().T
    return out


def test_asymmetric_two_qubit_kraus_both_orders():
    from invariantq_amd.core.binding import NativeDensityMatrix
    n = 5
    rng = np.random.default_rng(41)
    rho = random_rho(rng, 1 << n)
    # amplitude damping on the first wire, phase damping on the second: no
    # symmetry between the wires, so a swapped pair gives a different rho
    ks = [np.kron(a, b) for a in gates.amplitude_damping(0.3) for b in gates.phase_damping(0.2)]
    refs = {}
    for qs in ([4, 1], [1, 4], [0, 3]):
        print(f"2q Kraus on {qs}", flush=True)
        dm = NativeDensityMatrix(n)
        dm.set_matrix(rho)
        dm.apply_2q_kraus(qs[0], qs[1], ks)
        refs[tuple(qs)] = ref = kraus_dense(n, qs, ks, rho)
        assert np.max(np.abs(dm.get_matrix() - ref)) < 1e-14
        assert abs(dm.trace() - 1.0) < 1e-13
    assert np.max(np.abs(refs[(4, 1)] - refs[(1, 4)])) > 1e-3


@pytest.mark.parametrize("fuse", [False, True])
def test_simulator_new_channels_against_brute_force(fuse):
    from invariantq_amd import simulator
    n = 3
    ops = [("gate", "h", [0], None), ("gate", "cx", [0, 1], None), ("gate", "ry", [2], gates.ry(0.7)),
           ("noise", "thermal_relaxation", [0, 2], [40e-6, 60e-6, 2e-6]),
           ("noise", "depolarizing_2q", [2, 1], [0.3]),
           ("noise", "no_such_channel", [1], []),
           ("gate", "cx", [2, 0], None),
           ("noise", "thermal", [1], [])]
    res = simulator.DensityMatrixSimulator(fuse=fuse).run(n, ops, shots=64)
    rho = np.zeros((1 << n, 1 << n), dtype=complex)
    rho[0, 0] = 1.0
    for kind, name, qs, arg in ops:
        if kind == "gate":
            U = orc.full_operator(n, qs, arg if arg is not None else gates.gate_matrix(name))
            rho = U @ rho @ U.conj().T
        else:
            ks = gates.build_kraus_channel(name, arg)
            if ks[0].shape == (4, 4):
                rho = kraus_dense(n, qs, ks, rho)
            else:
                for q in qs:
                    rho = kraus_dense(n, [q], ks, rho)
    assert np.max(np.abs(res.density_matrix - rho)) < 1e-13
    assert abs(res.metadata["trace"] - 1.0) < 1e-13


def test_noisy_circuit_14_qubits_op_by_op_and_fused():
    from invariantq_amd.core.binding import NativeDensityMatrix
    n = 14
    ops = workloads.noisy_dm_circuit(n, depth=10, seed=77)
    print(f"noisy density matrix n={n}: {len(ops)} ops", flush=True)
    dm = NativeDensityMatrix(n)
    for op in ops:
        if op.kind == "kraus":
            dm.apply_1q_kraus(op.qubits[0], op.kraus)
        elif len(op.qubits) == 1:
            dm.apply_1q_unitary(op.qubits[0], op.matrix)
        else:
            dm.apply_2q_unitary(op.qubits[0], op.qubits[1], op.matrix)
    trace, purity, z = dm.trace(), dm.purity(), np.asarray(dm.expval_z())
    assert abs(trace - 1.0) < 1e-12
    assert 0.0 < purity <= 1.0 + 1e-12
    # the channels are not unital on their own (amplitude damping), and the
    # noise is weak: the state stays far from maximally mixed
    assert purity > 1.0 / (1 << n)
    rho = dm.get_matrix()
    rng = np.random.default_rng(14)
    r = rng.integers(0, 1 << n, 200000)
    c = rng.integers(0, 1 << n, 200000)
    assert np.max(np.abs(rho[r, c] - np.conj(rho[c, r]))) < 1e-15
    diag = np.real(np.diag(rho))
    assert diag.min() > -1e-15 and abs(diag.sum() - trace) < 1e-12
    idx = np.arange(1 << n)
    z_diag = np.array([np.sum(diag * (1 - 2 * ((idx >> q) & 1))) for q in range(n)])
    assert np.max(np.abs(z_diag - z)) < 1e-12
    del rho
    print("fused superoperators", flush=True)
    fused = fusion.fuse_density_matrix(ops, 2)
    assert len(fused) < len(ops) // 5
    dm.reset()
    for f in fused:
        dm.apply_superop(f.qubits, f.matrix)
    assert abs(dm.trace() - trace) < 1e-12
    assert abs(dm.purity() - purity) < 1e-12
    assert np.max(np.abs(np.asarray(dm.expval_z()) - z)) < 1e-12
