"""Seeded random sequences of 1- to 5-qubit operations (Haar unitaries,
diagonal phases, CX) on random target sets against the oracle's restatement
of the reference's statevector.cpp:26-152, at sizes on both sides of the
engine's kernel choices: the lane-exchange forms and the streaming hint from
24 qubits, the matrix-core pass from 9, LDS tiles from k + 6.  Every kernel
path is reached with targets inside and outside a 128-byte line.
Tolerance: 1e-11 absolute for complex128, 2e-5 for complex64."""

import numpy as np
import pytest

from oracle.oracle import haar_unitary

pytestmark = pytest.mark.gpu

CX = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], dtype=complex)


def _ops(n, rng, count):
    ops = [([q], haar_unitary(rng, 2)) for q in range(n)]  # every amplitude non-zero
    for _ in range(count):
        k = int(rng.choice([1, 1, 2, 2, 2, 3, 4, 5])) if n >= 5 else int(rng.integers(1, 3))
        k = min(k, n)
        if rng.random() < 0.5:  # at least one target among qubits 0-5
            low = int(rng.integers(0, min(6, n)))
            rest = [int(q) for q in rng.choice([q for q in range(n) if q != low], size=k - 1, replace=False)]
            qs = [low] + rest
            rng.shuffle(qs)
        else:
            qs = [int(q) for q in rng.choice(n, size=k, replace=False)]
        r = rng.random()
        if r < 0.15:
            m = np.diag(np.exp(1j * rng.uniform(0, 2 * np.pi, 1 << k)))
        elif r < 0.25 and k == 2:
            m = CX
        else:
            m = haar_unitary(rng, 1 << k)
        ops.append((list(qs), m))
    return ops


def _apply(t, ops):
    for qs, m in ops:
        if len(qs) == 1:
            t.apply_1q(qs[0], m)
        elif len(qs) == 2:
            t.apply_2q(qs[0], qs[1], m)
        else:
            t.apply_fused(qs, m)


@pytest.mark.parametrize("n", [6, 7, 13, 23, 24])
@pytest.mark.parametrize("precision", ["complex128", "complex64"])
def test_random_operations_match_oracle(oracle_lib, n, precision):
    from invariantq_amd.core import binding
    rng = np.random.default_rng(1000 * n + (precision == "complex64"))
    ops = _ops(n, rng, 40 if n < 20 else 24)
    print(f"random ops n={n} {precision}: {len(ops)} ops", flush=True)
    sv = binding.NativeStateVector(n, precision=precision)
    ref = oracle_lib.statevector(n)
    _apply(sv, ops)
    _apply(ref, ops)
    a, b = sv.get_amplitudes(), ref.get_amplitudes()
    tol = 1e-11 if precision == "complex128" else 2e-5
    assert np.max(np.abs(a - b)) < tol
