"""Host transfers of the statevector (reference core/src/api.cpp:74-90):
set_amplitudes, get_amplitudes and get_probabilities move the state in chunks
of 2^22 amplitudes through two pinned host slots.  States of 23-25 qubits
span 2-8 chunks (both slots, reused); a round trip must return the stored
values bit for bit and the probabilities must be |a|^2 of them."""

import numpy as np
import pytest

from invariantq_amd.core import binding as B

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,prec", [(12, "complex128"), (23, "complex128"), (25, "complex128"), (24, "complex64")])
def test_round_trip_and_probabilities(n, prec):
    print(f"transfers n={n} {prec}", flush=True)
    rng = np.random.default_rng(n)
    a = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    a /= np.linalg.norm(a)
    sv = B.NativeStateVector(n, prec)
    sv.set_amplitudes(a)
    stored = a.astype(np.complex64).astype(np.complex128) if prec == "complex64" else a
    got = sv.get_amplitudes()
    assert np.array_equal(got, stored)
    p = sv.get_probabilities()
    ref = stored.real ** 2 + stored.imag ** 2
    assert np.allclose(p, ref, rtol=1e-15, atol=0.0)
    # a second export after a gate reuses the pinned slots
    h = np.array([[1, 1], [1, -1]]) / np.sqrt(2)
    sv.apply_1q(n - 1, h)
    got2 = sv.get_amplitudes()
    half = 1 << (n - 1)
    lo, hi = stored[:half], stored[half:]
    want = np.concatenate([(lo + hi) / np.sqrt(2), (lo - hi) / np.sqrt(2)])
    assert np.max(np.abs(got2 - want)) < (1e-6 if prec == "complex64" else 1e-14)


@pytest.mark.parametrize("n,prec", [(9, "complex128"), (23, "complex128"), (23, "complex64")])
def test_split_and_interleaved_transfers_agree(n, prec):
    """The reference's split-array entry points (c_api.h:19-20, what the
    reference binding calls) against the interleaved ones the package binding
    uses: the same values bit for bit both ways, statevector and density
    matrix."""
    print(f"split vs interleaved n={n} {prec}", flush=True)
    lib = B.lib()
    rng = np.random.default_rng(n + 7)
    a = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    sv = B.NativeStateVector(n, prec)
    re, im = np.ascontiguousarray(a.real), np.ascontiguousarray(a.imag)
    lib.iq_statevector_set_amplitudes(sv._ptr, B._dp(re), B._dp(im))
    B.check_error()
    r2, i2 = np.empty(1 << n), np.empty(1 << n)
    lib.iq_statevector_get_amplitudes(sv._ptr, B._dp(r2), B._dp(i2))
    B.check_error()
    got = sv.get_amplitudes()  # interleaved
    assert np.array_equal(got.real, r2) and np.array_equal(got.imag, i2)
    sv2 = B.NativeStateVector(n, prec)
    sv2.set_amplitudes(a)  # interleaved
    lib.iq_statevector_get_amplitudes(sv2._ptr, B._dp(r2), B._dp(i2))
    B.check_error()
    assert np.array_equal(got.real, r2) and np.array_equal(got.imag, i2)
    if n <= 12:
        nq = n // 2
        rho = rng.standard_normal((1 << nq, 1 << nq)) + 1j * rng.standard_normal((1 << nq, 1 << nq))
        dm = B.NativeDensityMatrix(nq)
        rr, ri = np.ascontiguousarray(rho.real.reshape(-1)), np.ascontiguousarray(rho.imag.reshape(-1))
        lib.iq_density_matrix_set_elements(dm._ptr, B._dp(rr), B._dp(ri))
        B.check_error()
        assert np.array_equal(dm.get_matrix(), rho)
        dm.set_matrix(rho.T.copy())
        lib.iq_density_matrix_get_elements(dm._ptr, B._dp(rr), B._dp(ri))
        B.check_error()
        assert np.array_equal(rr.reshape(rho.shape) + 1j * ri.reshape(rho.shape), rho.T)
