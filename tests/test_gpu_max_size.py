"""The largest states one MI355X holds: 33 qubits complex128 and 34 qubits
complex64 (128 GiB each), whose indices run past bit 32.  A layer of one-qubit
gates makes a product state, whose per-qubit probabilities and Pauli
expectations have closed forms; 5-qubit blocks on high and low qubits followed
by their inverses must restore them; at 33 qubits the samples' bit frequencies
must follow the per-qubit probabilities."""

import numpy as np
import pytest

from invariantq_amd import workloads

pytestmark = pytest.mark.gpu


def _single(u):
    """<Z>, <X>, <Y> and P(0) of u|0>."""
    a, b = u[0, 0], u[1, 0]
    return (abs(a) ** 2 - abs(b) ** 2, 2 * (np.conj(a) * b).real, 2 * (np.conj(a) * b).imag, abs(a) ** 2)


@pytest.mark.parametrize("n,precision", [(33, "complex128"), (34, "complex64")])
def test_largest_states(n, precision):
    from invariantq_amd.core import binding
    tol = 1e-11 if precision == "complex128" else 2e-5
    print(f"largest state: n={n} {precision} ({(1 << n) * (16 if precision == 'complex128' else 8) >> 30} GiB)",
          flush=True)
    sv = binding.NativeStateVector(n, precision=precision)
    layer = workloads.u3_layer(n, seed=33)
    for op in layer:
        sv.apply_1q(op.qubits[0], op.matrix)
    z, x, y, p0 = zip(*(_single(op.matrix) for op in layer))
    probe = [0, 5, 31, 32, n - 1]
    for q in probe:
        assert abs(sv.measure_prob(q, 0) - p0[q]) < tol

    # ZZ chain pieces, X on qubits across every tile pass, one XY and one Z
    terms = [(1.0, "ZZ", [q, q + 1]) for q in (0, 15, 30, 31, n - 2)]
    terms += [(0.5, "X", [q]) for q in (1, 7, 13, 20, 26, 31, 32, n - 1)]
    terms += [(0.3, "XY", [3, n - 1]), (-0.7, "Z", [32])]

    def exact():
        e = 0.0
        for c, ops, qs in terms:
            v = 1.0
            for o, q in zip(ops, qs):
                v *= {"Z": z, "X": x, "Y": y}[o][q]
            e += c * v
        return e

    # The reference drops the products of amplitudes with |psi_i|^2 < 1e-18,
    # and a product state of this size has many: both sums keep that skip and
    # agree with each other to rounding (imaginary part included); the closed
    # form (no skip) differs by the dropped products (about 1e-8 here), far
    # below what a wrong index or phase would give.
    e_ref = exact()
    tiled = sv.expectation_value_pauli_sum(terms)
    per_mask, _ = sv.expectation_value_pauli_sum(terms, per_term=True)
    assert abs(tiled - per_mask) < tol
    assert abs(per_mask - e_ref) < max(tol, 1e-6)

    rng = np.random.default_rng(34)
    sets = ([32, 1, 17, 28, 9], [0, 2, 4, 31, n - 5], [n - 1, n - 2, n - 3, 12, 6])
    blocks = [(list(qs), workloads.haar_unitary(rng, 32)) for qs in sets]
    for qs, m in blocks:
        sv.apply_fused(qs, m)
    assert abs(sv.norm_squared() - 1.0) < tol
    for qs, m in reversed(blocks):
        sv.apply_fused(qs, m.conj().T)
    for q in probe:
        assert abs(sv.measure_prob(q, 0) - p0[q]) < tol
    assert abs(sv.expectation_value_pauli_sum(terms) - tiled) < tol

    if precision == "complex128":
        # few shots: per-chunk searches after one read of the state
        shots = 2000
        idx = sv.sample(shots, seed=5)
        assert idx.max() < (1 << n)
        for q in probe:
            freq = np.mean((idx >> np.uint64(q)) & np.uint64(1))
            sigma = np.sqrt(p0[q] * (1 - p0[q]) / shots)
            assert abs(freq - (1 - p0[q])) < 5 * sigma + 1e-3
        # more uniforms than twice the 2^21 chunks: the written CDF (64 GiB)
        # beside the state and the binary search over indices past 2^32;
        # the first 4096 agree with the per-chunk search on the same uniforms
        u = np.random.default_rng(35).random((1 << 22) + 4096)
        print("written-CDF sampling at 33 qubits", flush=True)
        many = sv.sample_uniforms(u)
        few = sv.sample_uniforms(u[:4096])
        assert np.array_equal(many[:4096], few)
        assert many.max() < (1 << n) and many.max() >= (1 << 32)
        for q in probe:
            freq = np.mean((many >> np.uint64(q)) & np.uint64(1))
            sigma = np.sqrt(p0[q] * (1 - p0[q]) / many.size)
            assert abs(freq - (1 - p0[q])) < 5 * sigma + 1e-3
    del sv
