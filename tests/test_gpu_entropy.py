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


def _sparse_state(n, m, rng):
    """m random basis states with random complex weights (normalised)."""
    idx = rng.choice(1 << n, size=m, replace=False)
    w = rng.normal(size=m) + 1j * rng.normal(size=m)
    psi = np.zeros(1 << n, dtype=complex)
    psi[idx] = w / np.linalg.norm(w)
    return psi, idx


def _exact_entropy(psi, n, keep):
    """Entropy of rho_A by dense diagonalisation of its non-zero block."""
    nz = np.nonzero(psi)[0]
    kept = [sum(((int(i) >> q) & 1) << b for b, q in enumerate(keep)) for i in nz]
    env = [int(i) & ~sum(1 << q for q in keep) for i in nz]
    rows = sorted(set(kept))
    envs = sorted(set(env))
    m = np.zeros((len(rows), len(envs)), dtype=complex)
    for i, r, e in zip(nz, kept, env):
        m[rows.index(r), envs.index(e)] = psi[i]
    lam = np.linalg.eigvalsh(m @ m.conj().T)
    lam = lam[lam > 1e-14]
    return float(-np.sum(lam * np.log2(lam)))


@pytest.mark.parametrize("n,k", [(14, 13), (17, 15)])
def test_entropy_beyond_12_kept_qubits(n, k):
    """13 and 15 kept qubits (rho_A of 2^26 and 2^30 entries): a state on four
    basis states has a 4x4 non-zero block, which the capped Jacobi
    diagonalises well inside its 100 rotations, so the entropy is the exact one."""
    from invariantq_amd.core.binding import NativeStateVector
    rng = np.random.default_rng(n * 100 + k)
    psi, _ = _sparse_state(n, 4, rng)
    keep = [int(q) for q in rng.permutation(n)[:k]]
    print(f"entropy n={n} k={k}", flush=True)
    sv = NativeStateVector(n)
    sv.set_amplitudes(psi)
    h = sv.von_neumann_entropy(keep)
    assert abs(h - _exact_entropy(psi, n, keep)) < 1e-9


def test_reduced_matrix_13_qubits_matches_oracle(oracle_lib):
    n, keep = 14, [13, 0, 12, 1, 11, 2, 10, 3, 9, 4, 8, 5, 7]
    sv, ref = states(oracle_lib, n, seed=21)
    print(f"reduced matrix n={n} k={len(keep)}", flush=True)
    got = sv.reduced_density_matrix(keep)
    want = ref.partial_trace(keep)
    assert np.max(np.abs(got - want)) < 1e-13


def test_entropy_above_15_is_nan_not_a_value():
    import ctypes
    from invariantq_amd.core import binding
    lib = binding.lib()
    sv = binding.NativeStateVector(17)
    q = np.arange(16, dtype=np.uint64)
    lib.iq_clear_error()
    h = lib.iq_statevector_entropy(sv._ptr, q.ctypes.data_as(ctypes.POINTER(ctypes.c_size_t)), 16)
    assert np.isnan(h)
    assert lib.iq_last_error() == 5  # IQ_ERR_UNSUPPORTED
    lib.iq_clear_error()
    # an empty or full subsystem is 0, as in the reference
    assert lib.iq_statevector_entropy(sv._ptr, q.ctypes.data_as(ctypes.POINTER(ctypes.c_size_t)), 0) == 0.0


def test_reduced_matrix_blocked_kernel_off_diagonal_blocks(oracle_lib):
    """10 kept of 16 (a 1024 x 1024 rho_A: 16 x 16 blocks of 64, the blocked
    matrix-core kernel with its off-diagonal blocks and their conjugate
    transposes), kept qubits in scrambled order, against the oracle."""
    n, keep = 16, [9, 2, 15, 0, 7, 4, 12, 1, 10, 5]
    sv, ref = states(oracle_lib, n, seed=31)
    print(f"reduced matrix n={n} k={len(keep)} (blocked)", flush=True)
    got, want = sv.reduced_density_matrix(keep), ref.partial_trace(keep)
    assert np.max(np.abs(got - want)) < 1e-13
    assert np.max(np.abs(got - got.conj().T)) < 1e-15


def _entropy_both_ways(sv, keep):
    """The entropy with the Jacobi on the device (rho_A never leaves it) and
    on the host (rho_A downloaded), same state, same rho_A."""
    from invariantq_amd.core import binding
    try:
        binding.set_entropy_options(1)
        dev = sv.von_neumann_entropy(keep)
        binding.set_entropy_options(16)
        host = sv.von_neumann_entropy(keep)
    finally:
        binding.set_entropy_options(0)
    return dev, host


@pytest.mark.parametrize("case", ["brick", "deep", "plus", "ghz", "sparse", "complex64"])
def test_device_jacobi_matches_host_bit_for_bit(oracle_lib, case):
    """The device Jacobi (rows rotated and scanned on the GPU, pivots and
    moduli on the host) against the host's row-maximum Jacobi, which
    tests/jacobi_check.cpp pins bit for bit to the oracle's full scan: the
    same entropy to the last bit.  The states cover a low-rank rho_A (most
    rows rescanned per rotation), a near-maximally mixed one, rows of one
    repeated modulus (|+>^n: every candidate ties), zero rows (GHZ, sparse),
    and complex64 amplitudes gathered into FP64."""
    from invariantq_amd.core.binding import NativeStateVector
    rng = np.random.default_rng(77)
    if case in ("brick", "deep", "complex64"):
        n = 18
        sv, _ = states(oracle_lib, n, "complex64" if case == "complex64" else "complex128",
                       seed=5 if case == "brick" else 6)
        if case == "deep":
            for op in workloads.random_u3_cx(n, 14, seed=66):
                if len(op.qubits) == 1:
                    sv.apply_1q(op.qubits[0], op.matrix)
                else:
                    sv.apply_2q(op.qubits[0], op.qubits[1], op.matrix)
        keeps = [list(range(9)), [int(q) for q in rng.permutation(n)[:12]]]
    else:
        n = 16
        if case == "plus":
            psi = np.full(1 << n, 2.0 ** (-n / 2), dtype=complex)
        elif case == "ghz":
            psi = np.zeros(1 << n, dtype=complex)
            psi[0] = psi[-1] = np.sqrt(0.5)
        else:
            psi, _ = _sparse_state(n, 64, rng)
        sv = NativeStateVector(n)
        sv.set_amplitudes(psi)
        keeps = [[3, 0, 9, 12, 5, 14, 1, 7], list(range(12))]
    for keep in keeps:
        print(f"device Jacobi {case} n={n} k={len(keep)}", flush=True)
        dev, host = _entropy_both_ways(sv, keep)
        assert np.isfinite(dev)
        assert dev == host, (dev, host)


@pytest.mark.parametrize("n,keep,precision", [
    (20, [7], "complex128"), (20, [19, 0], "complex128"), (20, [3, 17, 9], "complex128"),
    (19, [2, 11, 5], "complex64"), (20, [0, 1, 2, 3], "complex128"), (20, [18, 2, 9, 0, 13], "complex128"),
    (18, [4, 0, 12, 7, 1, 15, 9], "complex128"), (20, list(range(9)), "complex128"),
    (20, [8, 1, 19, 3, 11, 0, 6, 15, 4], "complex64")])
def test_reduced_matrix_environment_chunks(oracle_lib, n, keep, precision):
    """Few kept qubits of a larger state, where the environment sum is split
    over many workgroups: 1-3 kept (lane per environment index), 4-5 kept
    (one wave per 16 x 16 tile, chunked), 7 and 9 kept (64 x 64 blocks in
    both block orders, chunked), against the oracle element by element;
    the matrix is exactly Hermitian and the entropy matches."""
    sv, ref = states(oracle_lib, n, precision, seed=n + len(keep))
    print(f"reduced matrix n={n} keep={keep} {precision}", flush=True)
    got, want = sv.reduced_density_matrix(keep), ref.partial_trace(keep)
    tol = 1e-13 if precision == "complex128" else 2e-6
    assert np.max(np.abs(got - want)) < tol
    assert np.array_equal(got, got.conj().T)
    assert abs(np.trace(got).real - 1.0) < (1e-12 if precision == "complex128" else 1e-5)
    assert abs(sv.von_neumann_entropy(keep) - ref.von_neumann_entropy(keep)) < (1e-9 if precision == "complex128" else 1e-4)


def test_reduced_matrix_14_kept_matrix_core_direct_loads(oracle_lib):
    """14 kept qubits of 18 (d = 16384): the product whose strips arrive by
    direct loads into LDS, in the three-product form, against the oracle
    element by element (rows compared in chunks), exactly Hermitian."""
    n, keep = 18, [13, 0, 17, 2, 9, 4, 15, 6, 11, 8, 1, 10, 3, 12]
    sv, ref = states(oracle_lib, n, seed=41)
    print(f"reduced matrix n={n} k={len(keep)} (direct loads)", flush=True)
    got = sv.reduced_density_matrix(keep)
    want = ref.partial_trace(keep)
    step = 1024
    for r in range(0, got.shape[0], step):
        assert np.max(np.abs(got[r:r + step] - want[r:r + step])) < 1e-13
    assert abs(np.trace(got).real - 1.0) < 1e-12
    for r in range(0, got.shape[0], step):
        assert np.array_equal(got[r:r + step], got[:, r:r + step].conj().T)
