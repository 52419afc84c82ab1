"""Gate lists run through the circuit engine's tile passes
(StateVectorSimulator(tiled=True), GpuCircuit.from_gates): final states
against the oracle's op-by-op state (core/src/statevector.cpp:26-87 for every
gate) and against the engine's own op-by-op path; samples from the copied
state against the oracle's."""

import numpy as np
import pytest

from invariantq_amd import gates, workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    from invariantq_amd import simulator
    return simulator


def _oracle_state(oracle_lib, n, ops):
    ref = oracle_lib.statevector(n)
    for op in ops:
        if len(op.qubits) == 1:
            ref.apply_1q(op.qubits[0], op.matrix)
        else:
            ref.apply_2q(op.qubits[0], op.qubits[1], op.matrix)
    return ref


@pytest.mark.parametrize("n,layers", [(16, 6), (20, 20), (23, 3)])
def test_brick_circuit_matches_oracle(S, oracle_lib, n, layers):
    print(f"tiled brick circuit n={n} layers={layers}", flush=True)
    ops = workloads.random_u3_cx(n, layers, seed=100 + n)
    sv, passes = S.StateVectorSimulator(tiled=True).evolve(n, ops)
    ref = _oracle_state(oracle_lib, n, ops)
    err = float(np.max(np.abs(sv.get_amplitudes() - ref.get_amplitudes())))
    assert err < 1e-12, err
    # every layer is one stage of runs and a CZ set (the last layer's trailing
    # H runs one more): at most three passes per stage below 25 qubits, not
    # one per gate
    assert 0 < passes <= 3 * (layers + 1) < len(ops)
    assert np.array_equal(sv.sample(4000, 77), ref.sample(4000, 77))


def test_matches_op_by_op_with_generic_gates(S):
    """CZ, CX on either order of qubits, a Haar two-qubit gate (a pass of its
    own) and one-qubit gates between, at 17 qubits."""
    n = 17
    rng = np.random.default_rng(3)
    ops = workloads.u3_layer(n, seed=9)
    pairs = [(0, 16), (16, 0), (3, 4), (12, 2), (15, 14), (7, 9)]
    for a, b in pairs:
        ops.append(workloads.Op("gate", [a, b], gates.CX, name="cx"))
        ops.append(workloads.Op("gate", [b, a], gates.CZ, name="cz"))
        ops.append(workloads.Op("gate", [a], workloads.haar_unitary(rng, 2), name="u"))
    ops.append(workloads.Op("gate", [5, 13], workloads.haar_unitary(rng, 4), name="u2"))
    ops.append(workloads.Op("gate", [13, 5], workloads.haar_unitary(rng, 4), name="u2"))
    ops += workloads.u3_layer(n, seed=10)
    tiled, _ = S.StateVectorSimulator(tiled=True).evolve(n, ops)
    plain, _ = S.StateVectorSimulator().evolve(n, ops)
    assert np.max(np.abs(tiled.get_amplitudes() - plain.get_amplitudes())) < 1e-12


def test_complex64_against_oracle(S, oracle_lib):
    n = 18
    ops = workloads.random_u3_cx(n, 8, seed=4)
    sv, _ = S.StateVectorSimulator(precision="complex64", tiled=True).evolve(n, ops)
    ref = _oracle_state(oracle_lib, n, ops)
    amps = sv.get_amplitudes()
    assert np.max(np.abs(amps - ref.get_amplitudes())) < 2e-5
    assert abs(np.vdot(ref.get_amplitudes(), amps)) ** 2 > 1 - 1e-6


def test_resets_and_small_states_take_the_op_path(S):
    n = 16
    ops = workloads.random_u3_cx(n, 2, seed=8)
    with_reset = ops[:20] + [("reset", 3)] + ops[20:]
    a, passes = S.StateVectorSimulator(tiled=True).evolve(n, with_reset)
    b, _ = S.StateVectorSimulator().evolve(n, with_reset)
    assert passes is None
    assert np.array_equal(a.get_amplitudes(), b.get_amplitudes())
    small = workloads.random_u3_cx(8, 3, seed=8)
    c, passes = S.StateVectorSimulator(tiled=True).evolve(8, small)
    assert passes is None


def test_run_result_and_copy_shape_check(S):
    from invariantq_amd import adjoint
    from invariantq_amd.core import binding
    n = 16
    ops = workloads.random_u3_cx(n, 3, seed=11)
    res_t = S.StateVectorSimulator(tiled=True).run(n, ops, shots=300, seed=5, entropy=False)
    res_p = S.StateVectorSimulator().run(n, ops, shots=300, seed=5, entropy=False)
    assert res_t.counts == res_p.counts and res_t.metadata["tiled"]
    assert np.max(np.abs(res_t.statevector - res_p.statevector)) < 1e-12
    c = adjoint.GpuCircuit.from_gates(n, [(op.qubits, op.matrix) for op in ops])
    c.run()
    with pytest.raises((ValueError, RuntimeError)):
        c.copy_state_to(binding.NativeStateVector(n + 1))
    with pytest.raises((ValueError, RuntimeError)):
        c.copy_state_to(binding.NativeStateVector(n, precision="complex64"))


def test_large_state_against_op_by_op(S):
    """28 qubits (4 GiB complex128): the tile passes (three per layer) against
    the op-by-op kernels, norm and amplitudes."""
    n = 28
    print(f"tiled brick circuit n={n} against op by op", flush=True)
    ops = workloads.random_u3_cx(n, 3, seed=28)
    tiled, passes = S.StateVectorSimulator(tiled=True).evolve(n, ops)
    a = tiled.get_amplitudes()
    del tiled
    plain, _ = S.StateVectorSimulator().evolve(n, ops)
    b = plain.get_amplitudes()
    del plain
    assert passes <= 3 * 4
    assert abs(float(np.vdot(a, a).real) - 1.0) < 1e-12
    assert float(np.max(np.abs(a - b))) < 1e-12
