"""Fused statevector and density-matrix runs on the GPU engine against the
oracle's op-by-op restatement (statevector.cpp / density_matrix.cpp), and
against the engine's own unfused runs."""

import numpy as np
import pytest

from invariantq_amd import fusion, gates, workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    from invariantq_amd import simulator
    return simulator


def oracle_sv(oracle_lib, n, ops):
    ref = oracle_lib.statevector(n)
    for op in ops:
        if len(op.qubits) == 1:
            ref.apply_1q(op.qubits[0], op.matrix)
        else:
            ref.apply_2q(op.qubits[0], op.qubits[1], op.matrix)
    return ref


def oracle_dm(oracle_lib, n, ops):
    dm = oracle_lib.density_matrix(n)
    for op in ops:
        if op.kind == "kraus":
            for q in op.qubits:
                dm.apply_1q_kraus(q, op.kraus)
        elif len(op.qubits) == 1:
            dm.apply_1q_unitary(op.qubits[0], op.matrix)
        else:
            dm.apply_2q_unitary(op.qubits[0], op.qubits[1], op.matrix)
    return dm


@pytest.mark.parametrize("n,layers", [(12, 8), (16, 10)])
def test_statevector_fused_matches_oracle(S, oracle_lib, n, layers):
    ops = workloads.random_u3_cx(n, layers, seed=n)
    print(f"fused statevector n={n} gates={len(ops)}", flush=True)
    res = S.StateVectorSimulator(fusion_width=5).run(n, ops, shots=2000, seed=77, entropy=False)
    ref = oracle_sv(oracle_lib, n, ops)
    assert res.metadata["fused_ops"] * 4 <= len(ops)
    assert np.max(np.abs(res.statevector - ref.get_amplitudes())) < 1e-12
    plain = S.StateVectorSimulator().run(n, ops, shots=2000, seed=77, entropy=False)
    assert np.max(np.abs(res.statevector - plain.statevector)) < 1e-12


def test_statevector_fused_complex64(S, oracle_lib):
    n = 14
    ops = workloads.random_u3_cx(n, 10, seed=3)
    res = S.StateVectorSimulator(precision="complex64", fusion_width=5).run(n, ops, shots=10, entropy=False)
    ref = oracle_sv(oracle_lib, n, ops).get_amplitudes()
    assert np.max(np.abs(res.statevector - ref)) < 2e-5
    assert abs(np.vdot(ref, res.statevector)) ** 2 > 1 - 1e-6


@pytest.mark.parametrize("width", [3, 4])
def test_statevector_fused_narrow_widths(S, oracle_lib, width):
    n = 11
    ops = workloads.random_u3_cx(n, 6, seed=width)
    res = S.StateVectorSimulator(fusion_width=width).run(n, ops, shots=10, entropy=False)
    assert np.max(np.abs(res.statevector - oracle_sv(oracle_lib, n, ops).get_amplitudes())) < 1e-12


@pytest.mark.parametrize("n,depth", [(5, 4), (7, 3)])
def test_density_matrix_fused_matches_oracle(S, oracle_lib, n, depth):
    ops = workloads.noisy_dm_circuit(n, depth=depth, seed=11, p_depol=0.02, gamma=0.03, lam=0.02)
    print(f"fused density matrix n={n} ops={len(ops)}", flush=True)
    res = S.DensityMatrixSimulator(fuse=True).run(n, ops, shots=100, seed=3)
    assert res.metadata["fused_ops"] * 3 <= len(ops)
    ref = oracle_dm(oracle_lib, n, ops)
    want = ref.get_matrix()
    assert np.max(np.abs(res.density_matrix - want)) < 1e-12
    assert abs(res.metadata["trace"] - ref.trace()) < 1e-12
    assert abs(res.metadata["purity"] - ref.purity()) < 1e-12
    plain = S.DensityMatrixSimulator().run(n, ops, shots=100, seed=3)
    assert np.max(np.abs(res.density_matrix - plain.density_matrix)) < 1e-12


def test_density_matrix_superop_entry_point(oracle_lib):
    from invariantq_amd.core.binding import NativeDensityMatrix
    n = 4
    ops = [("gate", "h", [1], gates.gate_matrix("h")), ("gate", "cx", [1, 3], gates.CX),
           ("noise", "amplitude_damping", [3], (0.2,)), ("reset", 1), ("gate", "u3", [0], gates.u3(0.3, 0.2, 0.1))]
    fused = fusion.fuse_density_matrix(ops, 2)
    dm = NativeDensityMatrix(n)
    for f in fused:
        dm.apply_superop(f.qubits, f.matrix)
    ref = oracle_lib.density_matrix(n)
    ref.apply_1q_unitary(1, gates.gate_matrix("h"))
    ref.apply_2q_unitary(1, 3, gates.CX)
    ref.apply_1q_kraus(3, gates.amplitude_damping(0.2))
    ref.reset_qubit(1)
    ref.apply_1q_unitary(0, gates.u3(0.3, 0.2, 0.1))
    assert np.max(np.abs(dm.get_matrix() - ref.get_matrix())) < 1e-13
    with pytest.raises(ValueError):
        dm.apply_superop([0, 1, 2], np.eye(64))
    with pytest.raises(ValueError):
        dm.apply_superop([2, 2], np.eye(16))
    with pytest.raises(ValueError):
        dm.apply_superop([4], np.eye(4))


def test_fused_sequence_matches_single_calls():
    """One-call sequence = the same gates through apply_fused, bit for bit (same
    kernels, same packed matrices); diagonal gates, every width, buffer regrowth."""
    from invariantq_amd.core.binding import NativeStateVector
    from oracle.oracle import haar_unitary
    rng = np.random.default_rng(21)
    n = 14
    seqs = []
    for count in (40, 7, 120):
        seq = []
        for _ in range(count):
            k = int(rng.integers(1, 6))
            qs = [int(x) for x in rng.choice(n, k, replace=False)]
            if rng.random() < 0.15:
                m = np.diag(np.exp(1j * rng.uniform(0, 2 * np.pi, 1 << k)))
            else:
                m = haar_unitary(rng, 1 << k)
            seq.append((qs, m))
        seqs.append(seq)
    a, b = NativeStateVector(n), NativeStateVector(n)
    for seq in seqs:
        print(f"sequence of {len(seq)} gates", flush=True)
        a.apply_fused_sequence(seq)
        for qs, m in seq:
            if len(qs) == 1:
                b.apply_1q(qs[0], m)
            elif len(qs) == 2:
                b.apply_2q(qs[0], qs[1], m)
            else:
                b.apply_fused(qs, m)
        assert np.array_equal(a.get_amplitudes(), b.get_amplitudes())


def test_fused_sequence_rejects_before_running():
    from invariantq_amd.core.binding import NativeStateVector
    n = 8
    sv = NativeStateVector(n)
    sv.apply_1q(0, gates.gate_matrix("h"))
    before = sv.get_amplitudes()
    bad = [([1], gates.gate_matrix("x")), ([2, 2], np.eye(4)), ([3], gates.gate_matrix("x"))]
    with pytest.raises(ValueError):
        sv.apply_fused_sequence(bad)
    with pytest.raises(ValueError):
        sv.apply_fused_sequence([([1], gates.gate_matrix("x")), ([9], gates.gate_matrix("x"))])
    assert np.array_equal(sv.get_amplitudes(), before)
