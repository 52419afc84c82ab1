"""The streaming (non-temporal) instantiations of the gate kernels produce the
same amplitudes, bit for bit, as the cached ones: only the cache hint
differs.  They are chosen at 24 qubits and above when no target sits inside a
128-byte line, so the check runs a 24-qubit state with targets on both sides
of that boundary, and pins the complex128 result to the oracle's
statevector.cpp restatement."""

import numpy as np
import pytest

from oracle.oracle import haar_unitary

pytestmark = pytest.mark.gpu


def _ops(n, seed):
    rng = np.random.default_rng(seed)
    ops = []
    for t in (0, 2, 3, 5, 17, 23):  # below and above the line bits (3 for complex128, 4 for complex64)
        ops.append(([t], haar_unitary(rng, 2)))
    for qs in ([3, 20], [4, 11], [1, 9], [22, 6]):
        ops.append((qs, haar_unitary(rng, 4)))
    # 3- and 4-qubit blocks: the direct kernels when the lowest target is at or
    # above the line bits, the tiled one otherwise
    for qs in ([5, 10, 19], [4, 8, 15, 21], [1, 6, 12, 20]):
        ops.append((qs, haar_unitary(rng, 1 << len(qs))))
    for qs in ([3, 7, 12, 18, 23], [0, 1, 2, 9, 14]):
        ops.append((qs, haar_unitary(rng, 32)))
    return ops


def _run(binding, n, precision, ops):
    sv = binding.NativeStateVector(n, precision=precision)
    for qs, m in ops:
        if len(qs) == 1:
            sv.apply_1q(qs[0], m)
        elif len(qs) == 2:
            sv.apply_2q(qs[0], qs[1], m)
        else:
            sv.apply_fused(qs, m)
    out = sv.get_amplitudes()
    del sv
    return out


@pytest.mark.parametrize("precision", ["complex128", "complex64"])
def test_streaming_kernels_bit_exact(oracle_lib, precision):
    from invariantq_amd.core import binding
    n = 24
    ops = _ops(n, 24)
    print(f"streaming vs cached, n={n} {precision}", flush=True)
    try:
        binding.set_mfma_options()  # per-target-set defaults: streaming where lines are whole
        streamed = _run(binding, n, precision, ops)
        binding.set_mfma_options(nontemporal=0)  # cached accesses everywhere
        cached = _run(binding, n, precision, ops)
    finally:
        binding.set_mfma_options()
    assert np.array_equal(streamed, cached)
    if precision == "complex128":
        ref = oracle_lib.statevector(n)
        for qs, m in ops:
            if len(qs) == 1:
                ref.apply_1q(qs[0], m)
            elif len(qs) == 2:
                ref.apply_2q(qs[0], qs[1], m)
            else:
                ref.apply_fused(qs, m)
        assert np.max(np.abs(streamed - ref.get_amplitudes())) < 1e-12


@pytest.mark.parametrize("precision", ["complex128", "complex64"])
def test_streaming_diagonal_and_whole_state_passes_bit_exact(oracle_lib, precision):
    """From 2^26 amplitudes the diagonal-operator kernel and the whole-state
    passes (basis fill, normalisation) stream too: same bits as cached."""
    from invariantq_amd.core import binding
    n = 26
    rng = np.random.default_rng(26)
    ph = rng.uniform(0, 2 * np.pi, 6)
    ops = [([7], haar_unitary(rng, 2)),
           ([5], np.diag(np.exp(1j * ph[:2]))),
           ([3, 20], np.diag(np.exp(1j * ph[2:6]))),
           ([0, 25], haar_unitary(rng, 4))]
    print(f"streaming diagonal / whole-state passes, n={n} {precision}", flush=True)

    def run():
        sv = binding.NativeStateVector(n, precision=precision)
        for qs, m in ops:
            if len(qs) == 1:
                sv.apply_1q(qs[0], m)
            else:
                sv.apply_2q(qs[0], qs[1], m)
        sv.normalize()
        out = sv.get_amplitudes()
        del sv
        return out

    try:
        binding.set_mfma_options()
        streamed = run()
        binding.set_mfma_options(nontemporal=0)
        cached = run()
    finally:
        binding.set_mfma_options()
    assert np.array_equal(streamed, cached)
    if precision == "complex128":
        ref = oracle_lib.statevector(n)
        for qs, m in ops:
            if len(qs) == 1:
                ref.apply_1q(qs[0], m)
            else:
                ref.apply_2q(qs[0], qs[1], m)
        assert np.max(np.abs(streamed - ref.get_amplitudes())) < 1e-12
