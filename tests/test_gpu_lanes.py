"""One- and two-qubit gates on qubits 0-5 against the oracle's restatement
of the reference's statevector.cpp:26-87: the direct kernels on states up to
the 256 MiB Infinity Cache (8, 12, 24 qubits), the lane-exchange kernels
(a target inside a 128-byte line) and streaming beyond it (25 qubits
complex128, 26 complex64), with pairs in both orders and pairs that leave
the lane range (strided kernel).  Tolerance: 1e-12 absolute for
complex128, 2e-6 for complex64."""

import numpy as np
import pytest

from oracle.oracle import haar_unitary

pytestmark = pytest.mark.gpu

CX = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], dtype=complex)
CZ = np.diag([1, 1, 1, -1]).astype(complex)


def _ops(n, rng):
    ops = [([q], haar_unitary(rng, 2)) for q in range(n)]  # every amplitude non-zero
    ops += [([t], haar_unitary(rng, 2)) for t in range(6)]
    for pair in ([0, 1], [1, 0], [2, 5], [5, 3], [0, 4], [3, 2], [1, 7 % n], [n - 1, 2], [4, 5]):
        if len(set(pair)) == 2 and max(pair) < n:
            ops.append((pair, haar_unitary(rng, 4)))
    ops += [([0, 1], CX), ([2, 0], CX), ([1, 3], CZ)]
    return ops


@pytest.mark.parametrize("n, precision", [(8, "complex128"), (12, "complex128"), (24, "complex128"), (25, "complex128"),
                                          (8, "complex64"), (12, "complex64"), (24, "complex64"), (26, "complex64")])
def test_low_target_gates_match_oracle(oracle_lib, n, precision):
    from invariantq_amd.core import binding
    rng = np.random.default_rng(60 + n)
    ops = _ops(n, rng)
    print(f"low-target gates n={n} {precision}", flush=True)
    sv = binding.NativeStateVector(n, precision=precision)
    ref = oracle_lib.statevector(n)
    for qs, m in ops:
        for t in (sv, ref):
            if len(qs) == 1:
                t.apply_1q(qs[0], m)
            else:
                t.apply_2q(qs[0], qs[1], m)
    a, b = sv.get_amplitudes(), ref.get_amplitudes()
    tol = 1e-12 if precision == "complex128" else 2e-6
    assert np.max(np.abs(a - b)) < tol
