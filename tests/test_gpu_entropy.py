"""Partial trace and entanglement entropy on the GPU against the oracle's
restatement of core/src/statevector.cpp:293-410: the reduced matrix element
by element (matrix-core path for >= 4 kept and >= 2 traced qubits, the
one-thread-per-element kernel otherwise), the entropy to 1e-9."""

import numpy as np
import pytest

from invariantq_amd import workloads

pytestmark = pytest.mark.gpu


def states(oracle_lib, n, precision="complex128", seed=8):
    from invariantq_amd.core.binding import NativeStateVector
    ops = workloads.random_u3_cx(n, 6, seed=seed)
    sv, ref = NativeStateVector(n, precision=precision), oracle_lib.statevector(n)
    for op in ops:
        for s in (sv, ref):
            if len(op.qubits) == 1:
                s.apply_1q(op.qubits[0], op.matrix)
            else:
                s.apply_2q(op.qubits[0], op.qubits[1], op.matrix)
    return sv, ref


@pytest.mark.parametrize("keep", [[0, 1, 2, 3, 4, 5], [3, 0, 7, 1], [11, 10, 9, 8, 7, 6], [5], [2, 9],
                                  list(range(11)), [1, 0, 2, 3, 4, 5, 6, 7]])
def test_reduced_matrix_matches_oracle(oracle_lib, keep):
    n = 12
    sv, ref = states(oracle_lib, n)
    print(f"reduced matrix keep={keep}", flush=True)
    got, want = sv.reduced_density_matrix(keep), ref.partial_trace(keep)
    assert got.shape == want.shape
    assert np.max(np.abs(got - want)) < 1e-13
    assert abs(np.trace(got).real - 1.0) < 1e-12
    assert np.max(np.abs(got - got.conj().T)) < 1e-15


def test_reduced_matrix_complex64(oracle_lib):
    n = 12
    sv, ref = states(oracle_lib, n, "complex64")
    for keep in ([0, 1, 2, 3, 4], [6, 2, 9, 4]):
        assert np.max(np.abs(sv.reduced_density_matrix(keep) - ref.partial_trace(keep))) < 2e-6


@pytest.mark.parametrize("keep", [list(range(4)), list(range(8)), [9, 3, 12, 0, 5, 14]])
def test_entropy_matches_oracle(oracle_lib, keep):
    n = 16
    sv, ref = states(oracle_lib, n, seed=13)
    print(f"entropy keep={keep}", flush=True)
    assert abs(sv.von_neumann_entropy(keep) - ref.von_neumann_entropy(keep)) < 1e-9


def test_reduced_matrix_errors(oracle_lib):
    from invariantq_amd.core.binding import NativeStateVector
    sv = NativeStateVector(6)
    with pytest.raises(ValueError):
        sv.reduced_density_matrix([1, 1])
    with pytest.raises(ValueError):
        sv.reduced_density_matrix([6])
