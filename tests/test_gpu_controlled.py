"""Controlled one-qubit gates (a two- or three-qubit matrix blockdiag(I, U),
U on one qubit where the others are 1) are applied to the half (quarter) of
the state whose control bits are set.
Checked against NumPy's dense 4x4 product (reference core/src/statevector.cpp:49-87)
and against the dense GPU kernel on the same state, for CX, CY, CRY and a
controlled U3, with the control first or second in the qubit list, controls
above and below bit 4 (below, the dense kernel runs), both precisions; the
amplitudes whose control bit is clear must come back bit for bit.  The same
for Toffoli and a doubly controlled U3 through apply_fused (k = 3, reference
core/src/statevector.cpp:89-152), the target at each list position."""

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


def _numpy_apply(a, n, qs, m):
    k = len(qs)
    psi = a.reshape([2] * n)  # axis j is qubit n - 1 - j
    ax = [n - 1 - q for q in qs]
    out = np.tensordot(m.reshape([2] * (2 * k)), psi, axes=(list(range(k, 2 * k)), ax))
    return np.moveaxis(out, list(range(k)), ax).reshape(-1)


def _controlled_k(u, k, tpos):
    """blockdiag(I, U) on k qubits, U on list position tpos (MSB-first local index)."""
    d = 1 << k
    tb = 1 << (k - 1 - tpos)
    cb = (d - 1) & ~tb
    m = np.eye(d, dtype=np.complex128)
    idx = [cb, cb | tb]
    m[np.ix_(idx, idx)] = u
    return m


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
                    want = _numpy_apply(stored, n, [q0, q1], m)
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


@pytest.mark.parametrize("prec", ["complex128", "complex64"])
@pytest.mark.parametrize("n", [9, 12, 24])
def test_doubly_controlled_gates(n, prec):
    print(f"doubly controlled n={n} {prec}", flush=True)
    rng = np.random.default_rng(100 + n)
    a = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    a /= np.linalg.norm(a)
    tol = 1e-6 if prec == "complex64" else 1e-14
    stored = a.astype(np.complex64).astype(np.complex128) if prec == "complex64" else a
    sets = [(n - 1, 5, 0), (4, 6, n - 2), (2, 6, 8), (8, 1, 3)]
    for name in ("X", "U3"):
        for qs in sets:
            for tpos in range(3):
                m = _controlled_k(U1[name], 3, tpos)
                sv = B.NativeStateVector(n, prec)
                sv.set_amplitudes(a)
                sv.apply_fused(list(qs), m)
                got = sv.get_amplitudes()
                idx = np.arange(1 << n)
                ctrl = [q for j, q in enumerate(qs) if j != tpos]
                off = ((idx >> ctrl[0]) & 1 & (idx >> ctrl[1])) == 0
                if min(ctrl) >= 4:  # the controlled pass: the rest is not touched
                    assert np.array_equal(got[off], stored[off]), (name, qs, tpos)
                if n <= 12:
                    want = _numpy_apply(stored, n, list(qs), m)
                else:
                    B.set_gate_path("direct")
                    try:
                        ref = B.NativeStateVector(n, prec)
                        ref.set_amplitudes(a)
                        ref.apply_fused(list(qs), m)
                        want = ref.get_amplitudes()
                    finally:
                        B.set_gate_path("auto")
                assert np.max(np.abs(got - want)) < tol, (name, qs, tpos)
