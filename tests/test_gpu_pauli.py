"""Pauli sums on the GPU against the oracle's per-string restatement of the
reference's <psi|P|psi> (core/src/statevector.cpp:258-291), on 20 and 22
qubits, where sums over many flip masks run over tile passes.  Tolerance:
1e-12 absolute for complex128, 1e-6 relative for complex64 (FP32 amplitudes,
FP64 sums)."""

import numpy as np
import pytest

from invariantq_amd import workloads
from oracle.oracle import haar_unitary

pytestmark = pytest.mark.gpu


def _terms(n, rng):
    terms = list(workloads.tfim_terms(n))
    # strings whose flip masks span two tile passes, or none: Y and Z mixes
    for _ in range(12):
        k = int(rng.integers(1, 5))
        qs = [int(q) for q in rng.choice(n, size=k, replace=False)]
        ops = "".join(rng.choice(list("XYZ"), size=k))
        terms.append((float(rng.normal()), ops, qs))
    terms.append((0.3, "XX", [0, n - 1]))
    terms.append((-0.7, "YY", [3, n - 2]))
    terms.append((0.2, "ZIZ", [1, 5, 9]))  # identity characters act as I
    return terms


def _state(n, rng):
    # one-qubit layer, then overlapping 5-qubit blocks: every amplitude non-zero
    ops = [([q], haar_unitary(rng, 2)) for q in range(n)]
    for start in range(0, n - 4, 3):
        ops.append((list(range(start, start + 5)), haar_unitary(rng, 32)))
    return ops


def _apply(sv, ops):
    for qs, m in ops:
        if len(qs) == 1:
            sv.apply_1q(qs[0], m)
        else:
            sv.apply_fused(qs, m)


def _oracle_sum(ref, terms):
    total = 0.0 + 0.0j
    for c, o, qs in terms:
        total += c * ref.expectation_value_pauli(o, qs)
    return total


@pytest.mark.parametrize("n,precision", [(20, "complex128"), (22, "complex128"), (22, "complex64")])
def test_tiled_pauli_sum_matches_oracle(oracle_lib, n, precision):
    from invariantq_amd.core import binding
    rng = np.random.default_rng(n)
    ops = _state(n, rng)
    terms = _terms(n, rng)
    print(f"pauli sum n={n} {precision}, {len(terms)} strings", flush=True)
    sv = binding.NativeStateVector(n, precision=precision)
    _apply(sv, ops)
    tiled = sv.expectation_value_pauli_sum(terms)
    per_mask, vals = sv.expectation_value_pauli_sum(terms, per_term=True)
    ref = oracle_lib.statevector(n)
    _apply(ref, ops)
    exact = _oracle_sum(ref, terms)
    if precision == "complex128":
        assert abs(tiled - exact) < 1e-12
        assert abs(per_mask - exact) < 1e-12
        singles = np.array([ref.expectation_value_pauli(o, qs) for _, o, qs in terms])
        assert np.max(np.abs(vals - singles)) < 1e-12
    else:
        scale = sum(abs(c) for c, _, _ in terms)
        assert abs(tiled - exact) < 1e-6 * scale
        assert abs(tiled - per_mask) < 1e-6 * scale


def test_tiled_pauli_sum_keeps_the_reference_skip(oracle_lib):
    """Amplitudes with |psi_i|^2 < 1e-18 add nothing to <P> in the reference;
    a state with many of them (here scaled to 3e-10) separates a sum that
    keeps that skip from one that does not."""
    from invariantq_amd.core import binding
    n = 20
    rng = np.random.default_rng(7)
    ops = _state(n, rng)
    sv = binding.NativeStateVector(n)
    _apply(sv, ops)
    psi = sv.get_amplitudes()
    idx = np.arange(psi.size)
    small = ((idx >> 7) & 1) == 1  # half the amplitudes, in runs of 128
    psi[small] *= 3e-10 / np.abs(psi[small]).max()
    sv.set_amplitudes(psi)
    ref = oracle_lib.statevector(n)
    ref.set_amplitudes(psi)
    terms = _terms(n, rng)
    exact = _oracle_sum(ref, terms)
    tiled = sv.expectation_value_pauli_sum(terms)
    # the skipped products are far above the tolerance: check that they matter
    no_skip = 0.0
    for c, o, qs in terms:
        no_skip += c * np.vdot(psi, _pauli_apply(psi, o, qs)).real
    assert abs(no_skip - exact.real) > 1e-9
    # the skip drops conj(psi_j) psi_i but keeps its mirror: a non-real sum
    assert abs(exact.imag) > 1e-13
    assert abs(tiled - exact) < 1e-12


def _pauli_apply(psi, ops, qubits):
    """P |psi> for a Pauli string, by index arithmetic (NumPy)."""
    out = psi.copy()
    idx = np.arange(psi.size)
    for o, q in zip(ops, qubits):
        bit = (idx >> q) & 1
        if o == "X":
            out = out[idx ^ (1 << q)]
        elif o == "Y":
            out = out[idx ^ (1 << q)] * np.where(bit == 1, 1j, -1j)
        elif o == "Z":
            out = out * np.where(bit == 1, -1.0, 1.0)
    return out
