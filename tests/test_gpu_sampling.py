"""Sampling's two GPU searches against the exported CDF.

A shot's index is the first whose CDF value reaches the uniform, clamped to
dim - 1 (the lower_bound of reference core/src/statevector.cpp:248-254).  With
few shots each shot re-sums only the 4096-amplitude chunk it lands in; with
many, the CDF is written and searched.  Both must give exactly
searchsorted(cdf, u, "left") over the exported CDF, which must be monotone and
within rounding of the serial running sum, for uniforms placed on, just below and
just above CDF values at chunk and thread-segment boundaries, on flat
stretches of zero probability, at 0 and above the largest value; 26 qubits
spans several 4096-chunk groups of the offset scan.  Agreement of
the samples with the oracle's serial CDF is tested with the golden vectors and
the 20-qubit circuit."""

import numpy as np
import pytest

from invariantq_amd.core import binding as B

pytestmark = pytest.mark.gpu

CHUNK = 4096
SEGMENT = 16


def _random_state(n, seed, zero_frac=0.0):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    if zero_frac:
        a[rng.random(1 << n) < zero_frac] = 0.0
    return a / np.linalg.norm(a)


def _adversarial_uniforms(cdf, seed):
    rng = np.random.default_rng(seed)
    dim = cdf.size
    idx = {0, dim - 1}
    for b in range(0, dim, CHUNK):  # chunk edges
        idx.update({b - 1, b, b + 1})
    for b in range(0, min(dim, 4 * CHUNK), SEGMENT):  # thread-segment edges
        idx.update({b - 1, b})
    idx.update(rng.integers(0, dim, 64).tolist())
    idx = np.array(sorted(i for i in idx if 0 <= i < dim))
    vals = cdf[idx]
    u = np.concatenate([vals, np.nextafter(vals, 0.0), np.nextafter(vals, 2.0), rng.random(256),
                        [0.0, cdf[-1], np.nextafter(cdf[-1], 2.0), 1.5]])
    return u[(u >= 0.0)]


def _switch(n):
    """Most shots per call that use the per-chunk searches: twice the number
    of 4096-amplitude chunks."""
    return 2 * max(1, (1 << n) // CHUNK)


def _expected(cdf, u):
    return np.minimum(np.searchsorted(cdf, u, side="left"), cdf.size - 1).astype(np.uint64)


@pytest.mark.parametrize("prec", ["complex128", "complex64"])
@pytest.mark.parametrize("n,zero_frac", [(3, 0.0), (10, 0.5), (13, 0.0), (14, 0.9), (20, 0.0), (26, 0.0)])
def test_chunk_and_cdf_searches_match_exported_cdf(n, zero_frac, prec):
    print(f"sampling n={n} zeros={zero_frac} {prec}", flush=True)
    sv = B.NativeStateVector(n, prec)
    a = _random_state(n, 100 + n, zero_frac)
    sv.set_amplitudes(a)
    cdf = sv.get_cdf()
    assert np.all(np.diff(cdf) >= 0.0)
    stored = a.astype(np.complex64).astype(np.complex128) if prec == "complex64" else a
    # the serial running sum's own rounding grows with the length
    tol = max(1e-12, (1 << n) * 2.0 ** -53)
    assert np.max(np.abs(cdf - np.cumsum(stored.real ** 2 + stored.imag ** 2))) < tol
    u = _adversarial_uniforms(cdf, n)
    want = _expected(cdf, u)
    switch = _switch(n)
    # many shots in one call: the written CDF (padded past the switch)
    many = np.concatenate([u, np.full(switch + 1, 0.5)])
    assert np.array_equal(sv.sample_uniforms(many)[: u.size], want)
    # few shots per call: per-chunk searches
    for s in range(0, u.size, switch):
        part = u[s: s + switch]
        assert np.array_equal(sv.sample_uniforms(part), want[s: s + switch]), (n, s)


def test_flat_cdf_between_two_amplitudes():
    """GHZ-like state over 13 qubits: the CDF is 0.5 from index 0 to dim-2.
    u = 0.5 gives 0; the next double up gives dim - 1 from both searches."""
    n = 13
    dim = 1 << n
    sv = B.NativeStateVector(n)
    a = np.zeros(dim, dtype=np.complex128)
    a[0] = a[-1] = np.sqrt(0.5)
    sv.set_amplitudes(a)
    cdf = sv.get_cdf()
    u = np.array([0.0, 0.25, cdf[0], np.nextafter(cdf[0], 2.0), cdf[-1]])
    want = _expected(cdf, u)
    assert want[0] == 0 and want[2] == 0 and want[3] == dim - 1
    for x, w in zip(u, want):
        assert sv.sample_uniforms(np.array([x]))[0] == w
    assert np.array_equal(sv.sample_uniforms(np.concatenate([u, np.zeros(_switch(n))]))[: u.size], want)


def test_sample_and_tasks_agree_across_search_paths():
    """sample(shots, seed) with shots on both sides of the path switch equals
    the corresponding prefix of one many-shot call with the same seed."""
    n = 16
    sv = B.NativeStateVector(n)
    sv.set_amplitudes(_random_state(n, 7))
    switch = _switch(n)
    big = sv.sample(4 * switch, seed=99)
    for shots in (1, switch // 2, switch, switch + 1):
        assert np.array_equal(sv.sample(shots, seed=99), big[:shots]), shots
    got = sv.sample_tasks([5, 6], [switch // 4, switch // 4])
    assert np.array_equal(got[0], sv.sample(switch // 4, seed=5))
    assert np.array_equal(got[1], sv.sample(switch // 4, seed=6))
