"""Controlled one-qubit gates (a two-qubit matrix blockdiag(I, U) over one of
its qubits) are applied to the half of the state whose control bit is set.
Checked against NumPy's dense 4x4 product (reference core/src/statevector.cpp:49-87)
and against the dense GPU kernel on the same state, for CX, CY, CRY and a
controlled U3, with the control first or second in the qubit list, controls
above and below bit 4 (below, the dense kernel runs), both precisions; the
amplitudes whose control bit is clear must come back bit for bit."""

import numpy as np
import pytest

from invariantq_amd import gates
from invariantq_amd.core import binding as B

pytestmark = pytest.mark.gpu


def _controlled(u, control_first):
    m = np.eye(4, dtype=np.complex128)
    if control_first:  # local index 2 b(q0) + b(q1): control q0 -> rows 2, 3
        m[2:, 2:] = u
    else:  # control q1 -> rows 1, 3
        m[np.ix_([1, 3], [1, 3])] = u
    return m


def _numpy_apply(a, n, q0, q1, m):
    psi = a.reshape([2] * n)  # axis k is qubit n - 1 - k
    out = np.tensordot(m.reshape(2, 2, 2, 2), psi, axes=([2, 3], [n - 1 - q0, n - 1 - q1]))
    return np.moveaxis(out, [0, 1], [n - 1 - q0, n - 1 - q1]).reshape(-1)


U1 = {
    "X": np.array([[0, 1], [1, 0]], dtype=np.complex128),
    "Y": np.array([[0, -1j], [1j, 0]]),
    "RY": np.array([[np.cos(0.35), -np.sin(0.35)], [np.sin(0.35), np.cos(0.35)]], dtype=np.complex128),
    "U3": gates.u3(0.7, -1.2, 2.4),
}


@pytest.mark.parametrize("prec", ["complex128", "complex64"])
@pytest.mark.parametrize("n", [8, 12, 24])
def test_controlled_gates(n, prec):
    print(f"controlled n={n} {prec}", flush=True)
    rng = np.random.default_rng(n)
    a = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    a /= np.linalg.norm(a)
    tol = 1e-6 if prec == "complex64" else 1e-14
    stored = a.astype(np.complex64).astype(np.complex128) if prec == "complex64" else a
    pairs = [(n - 1, 0), (5, 2), (4, n - 2), (1, n - 1), (2, 3)]
    for name, u in U1.items():
        for c, t in pairs:
            for control_first in (True, False):
                q0, q1 = (c, t) if control_first else (t, c)
                m = _controlled(u, control_first)
                sv = B.NativeStateVector(n, prec)
                sv.set_amplitudes(a)
                sv.apply_2q(q0, q1, m)
                got = sv.get_amplitudes()
                idx = np.arange(1 << n)
                off = ((idx >> c) & 1) == 0
                assert np.array_equal(got[off], stored[off]), (name, c, t, control_first)
                if n <= 12:
                    want = _numpy_apply(stored, n, q0, q1, m)
                else:
                    B.set_gate_path("direct")
                    try:
                        ref = B.NativeStateVector(n, prec)
                        ref.set_amplitudes(a)
                        ref.apply_2q(q0, q1, m)
                        want = ref.get_amplitudes()
                    finally:
                        B.set_gate_path("auto")
                assert np.max(np.abs(got - want)) < tol, (name, c, t, control_first)
