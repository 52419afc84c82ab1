"""TEST INFRASTRUCTURE ONLY — the oracle's Python side.

Two things live here:

* ``CoreLib`` — a ctypes binding shaped like the reference binding
  (invariantq/core/binding.py:54-299, 302-476) that can drive any library
  exporting the reference C names: the oracle restatement ``oracle/liboracle.so``
  and, in the build container only, the reference core compiled from source
  (``oracle/_ref/libinvariantq.so``).
* NumPy restatements used as independent checks: brute-force dense operators,
  the tensor-axis gate application of the reference fallback
  (invariantq/core/binding.py:603-620), Haar unitaries, the TFIM observable and
  FP64 parameter-shift / finite-difference gradients.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this module, and only as the checker or the timed CPU baseline.  The product
(invariantq_amd) never imports it.
"""

from __future__ import annotations

import ctypes
import hashlib
import math
import os
import subprocess
from pathlib import Path
from typing import List, Sequence, Tuple

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_LIB = HERE / "liboracle.so"
REF_LIB = HERE / "_ref" / "libinvariantq.so"
REFERENCE_CORE = Path("/root/reference/core")

_dp = ctypes.POINTER(ctypes.c_double)
_sp = ctypes.POINTER(ctypes.c_size_t)
_u64p = ctypes.POINTER(ctypes.c_uint64)


def build(ref: bool = True) -> None:
    """Compile the restatement, and the reference core where its sources exist."""
    subprocess.run(["make", "-s", "-C", str(HERE), "all"], check=True)
    if ref and REFERENCE_CORE.is_dir():
        subprocess.run(["make", "-s", "-C", str(HERE), "ref"], check=True)


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _split(m) -> Tuple[np.ndarray, np.ndarray]:
    flat = np.asarray(m, dtype=np.complex128).reshape(-1)
    return (np.ascontiguousarray(flat.real, dtype=np.float64),
            np.ascontiguousarray(flat.imag, dtype=np.float64))


class CoreLib:
    """Binds the 26 statevector / density-matrix names of core/include/c_api.h:10-38."""

    def __init__(self, path: Path | str):
        self.path = Path(path)
        self.lib = ctypes.CDLL(str(self.path))
        L = self.lib
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        sig = {
            "iq_statevector_create": ([sz], vp),
            "iq_statevector_free": ([vp], None),
            "iq_statevector_reset": ([vp], None),
            "iq_statevector_apply_1q": ([vp, sz, _dp, _dp], None),
            "iq_statevector_apply_2q": ([vp, sz, sz, _dp, _dp], None),
            "iq_statevector_apply_fused": ([vp, _sp, sz, _dp, _dp], None),
            "iq_statevector_measure": ([vp, sz, ctypes.c_double], ctypes.c_int),
            "iq_statevector_reset_qubit": ([vp, sz, ctypes.c_double], None),
            "iq_statevector_measure_prob": ([vp, sz, ctypes.c_int], ctypes.c_double),
            "iq_statevector_get_amplitudes": ([vp, _dp, _dp], None),
            "iq_statevector_get_probabilities": ([vp, _dp], None),
            "iq_statevector_sample": ([vp, sz, ctypes.c_uint64, _u64p], None),
            "iq_statevector_entropy": ([vp, _sp, sz], ctypes.c_double),
            "iq_statevector_expval_pauli": ([vp, ctypes.c_char_p, _sp, sz, _dp, _dp], None),
            "iq_density_matrix_create": ([sz], vp),
            "iq_density_matrix_free": ([vp], None),
            "iq_density_matrix_reset": ([vp], None),
            "iq_density_matrix_apply_1q_unitary": ([vp, sz, _dp, _dp], None),
            "iq_density_matrix_apply_2q_unitary": ([vp, sz, sz, _dp, _dp], None),
            "iq_density_matrix_apply_kraus_1q": ([vp, sz, sz, _dp, _dp], None),
            "iq_density_matrix_get_elements": ([vp, _dp, _dp], None),
            "iq_density_matrix_trace": ([vp], ctypes.c_double),
            "iq_density_matrix_purity": ([vp], ctypes.c_double),
            "iq_density_matrix_entropy": ([vp], ctypes.c_double),
            "iq_density_matrix_measure": ([vp, sz, ctypes.c_double], ctypes.c_int),
            "iq_density_matrix_reset_qubit": ([vp, sz], None),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        self.has_helpers = hasattr(L, "iqo_statevector_set_amplitudes")
        if self.has_helpers:
            L.iqo_statevector_set_amplitudes.argtypes = [vp, _dp, _dp]
            L.iqo_statevector_set_amplitudes.restype = None
            L.iqo_statevector_norm_squared.argtypes = [vp]
            L.iqo_statevector_norm_squared.restype = ctypes.c_double
            L.iqo_statevector_normalize.argtypes = [vp]
            L.iqo_statevector_normalize.restype = None
            L.iqo_density_matrix_expval_z.argtypes = [vp, _dp]
            L.iqo_density_matrix_expval_z.restype = None
            L.iqo_density_matrix_set_elements.argtypes = [vp, _dp, _dp]
            L.iqo_density_matrix_set_elements.restype = None
            L.iqo_last_error.argtypes = []
            L.iqo_last_error.restype = ctypes.c_int
            L.iqo_mt19937_64_raw.argtypes = [ctypes.c_uint64, sz, _u64p]
            L.iqo_mt19937_64_raw.restype = None
            L.iqo_uniforms.argtypes = [ctypes.c_uint64, sz, _dp]
            L.iqo_uniforms.restype = None
            L.iqo_statevector_partial_trace.argtypes = [vp, _sp, sz, _dp, _dp]
            L.iqo_statevector_partial_trace.restype = None
            L.iqo_hermitian_eigenvalues.argtypes = [_dp, _dp, sz, ctypes.c_int, ctypes.c_double, _dp]
            L.iqo_hermitian_eigenvalues.restype = None

    def mt19937_64_raw(self, seed: int, count: int) -> np.ndarray:
        out = np.empty(count, dtype=np.uint64)
        self.lib.iqo_mt19937_64_raw(seed, count, out.ctypes.data_as(_u64p))
        return out

    def uniforms(self, seed: int, count: int) -> np.ndarray:
        out = np.empty(count, dtype=np.float64)
        self.lib.iqo_uniforms(seed, count, _dptr(out))
        return out

    def statevector(self, n: int) -> "CoreStateVector":
        return CoreStateVector(self, n)

    def density_matrix(self, n: int) -> "CoreDensityMatrix":
        return CoreDensityMatrix(self, n)


class CoreStateVector:
    """Method names follow NativeStateVector (invariantq/core/binding.py:302-395)."""

    def __init__(self, core: CoreLib, n: int):
        self.core, self.L = core, core.lib
        self.num_qubits, self.dim = n, 1 << n
        self._ptr = self.L.iq_statevector_create(n)
        if not self._ptr:
            raise MemoryError("oracle statevector allocation failed")

    def __del__(self):
        if getattr(self, "_ptr", None):
            self.L.iq_statevector_free(self._ptr)
            self._ptr = None

    def reset(self):
        self.L.iq_statevector_reset(self._ptr)

    def apply_1q(self, t: int, m):
        re, im = _split(m)
        self.L.iq_statevector_apply_1q(self._ptr, t, _dptr(re), _dptr(im))

    def apply_2q(self, q0: int, q1: int, m):
        re, im = _split(m)
        self.L.iq_statevector_apply_2q(self._ptr, q0, q1, _dptr(re), _dptr(im))

    def apply_fused(self, qubits: Sequence[int], m):
        re, im = _split(m)
        q = np.ascontiguousarray(qubits, dtype=np.uint64)
        self.L.iq_statevector_apply_fused(self._ptr, q.ctypes.data_as(_sp), len(qubits), _dptr(re), _dptr(im))

    def measure(self, t: int, u: float) -> int:
        return self.L.iq_statevector_measure(self._ptr, t, u)

    def reset_qubit(self, t: int, u: float = 0.0):
        self.L.iq_statevector_reset_qubit(self._ptr, t, u)

    def measure_prob(self, t: int, outcome: int) -> float:
        return self.L.iq_statevector_measure_prob(self._ptr, t, outcome)

    def get_amplitudes(self) -> np.ndarray:
        re = np.empty(self.dim)
        im = np.empty(self.dim)
        self.L.iq_statevector_get_amplitudes(self._ptr, _dptr(re), _dptr(im))
        return re + 1j * im

    def set_amplitudes(self, psi):
        re, im = _split(psi)
        self.L.iqo_statevector_set_amplitudes(self._ptr, _dptr(re), _dptr(im))

    def get_probabilities(self) -> np.ndarray:
        p = np.empty(self.dim)
        self.L.iq_statevector_get_probabilities(self._ptr, _dptr(p))
        return p

    def sample(self, shots: int, seed: int = 42) -> np.ndarray:
        out = np.empty(shots, dtype=np.uint64)
        self.L.iq_statevector_sample(self._ptr, shots, seed, out.ctypes.data_as(_u64p))
        return out

    def partial_trace(self, qubits: Sequence[int]) -> np.ndarray:
        """Reduced density matrix (statevector.cpp:293-330); oracle library only."""
        q = np.ascontiguousarray(qubits, dtype=np.uint64)
        d = 1 << q.size
        re, im = np.empty(d * d), np.empty(d * d)
        self.L.iqo_statevector_partial_trace(self._ptr, q.ctypes.data_as(_sp), q.size, _dptr(re), _dptr(im))
        return (re + 1j * im).reshape(d, d)

    def von_neumann_entropy(self, qubits: Sequence[int]) -> float:
        q = np.ascontiguousarray(qubits, dtype=np.uint64)
        return self.L.iq_statevector_entropy(self._ptr, q.ctypes.data_as(_sp), len(qubits))

    def expectation_value_pauli(self, ops: str, qubits: Sequence[int]) -> complex:
        q = np.ascontiguousarray(qubits, dtype=np.uint64)
        vr, vi = ctypes.c_double(0.0), ctypes.c_double(0.0)
        self.L.iq_statevector_expval_pauli(self._ptr, ops.encode("ascii"), q.ctypes.data_as(_sp), len(qubits),
                                           ctypes.byref(vr), ctypes.byref(vi))
        return complex(vr.value, vi.value)

    def norm_squared(self) -> float:
        return self.L.iqo_statevector_norm_squared(self._ptr)

    def normalize(self):
        """StateVector::normalize (statevector.cpp:215-224); oracle library only."""
        self.L.iqo_statevector_normalize(self._ptr)


class CoreDensityMatrix:
    """Method names follow NativeDensityMatrix (invariantq/core/binding.py:398-476)."""

    def __init__(self, core: CoreLib, n: int):
        self.core, self.L = core, core.lib
        self.num_qubits, self.dim = n, 1 << n
        self._ptr = self.L.iq_density_matrix_create(n)
        if not self._ptr:
            raise MemoryError("oracle density matrix allocation failed")

    def __del__(self):
        if getattr(self, "_ptr", None):
            self.L.iq_density_matrix_free(self._ptr)
            self._ptr = None

    def reset(self):
        self.L.iq_density_matrix_reset(self._ptr)

    def apply_1q_unitary(self, t: int, m):
        re, im = _split(m)
        self.L.iq_density_matrix_apply_1q_unitary(self._ptr, t, _dptr(re), _dptr(im))

    def apply_2q_unitary(self, q0: int, q1: int, m):
        re, im = _split(m)
        self.L.iq_density_matrix_apply_2q_unitary(self._ptr, q0, q1, _dptr(re), _dptr(im))

    def apply_1q_kraus(self, t: int, kraus: Sequence):
        stacked = np.stack([np.asarray(k, dtype=np.complex128).reshape(-1) for k in kraus])
        re, im = _split(stacked)
        self.L.iq_density_matrix_apply_kraus_1q(self._ptr, t, len(kraus), _dptr(re), _dptr(im))

    def get_matrix(self) -> np.ndarray:
        total = self.dim * self.dim
        re = np.empty(total)
        im = np.empty(total)
        self.L.iq_density_matrix_get_elements(self._ptr, _dptr(re), _dptr(im))
        return (re + 1j * im).reshape(self.dim, self.dim)

    def set_matrix(self, rho):
        re, im = _split(rho)
        self.L.iqo_density_matrix_set_elements(self._ptr, _dptr(re), _dptr(im))

    def trace(self) -> float:
        return self.L.iq_density_matrix_trace(self._ptr)

    def purity(self) -> float:
        return self.L.iq_density_matrix_purity(self._ptr)

    def von_neumann_entropy(self) -> float:
        return self.L.iq_density_matrix_entropy(self._ptr)

    def measure(self, t: int, u: float) -> int:
        return self.L.iq_density_matrix_measure(self._ptr, t, u)

    def reset_qubit(self, t: int):
        self.L.iq_density_matrix_reset_qubit(self._ptr, t)

    def expval_z(self) -> np.ndarray:
        out = np.empty(self.num_qubits)
        self.L.iqo_density_matrix_expval_z(self._ptr, _dptr(out))
        return out


_ORACLE = None


def oracle() -> CoreLib:
    """The restatement library (built by ``build()`` / ``make -C oracle``)."""
    global _ORACLE
    if _ORACLE is None:
        if not ORACLE_LIB.exists():
            raise FileNotFoundError(f"{ORACLE_LIB} is not built; run make -C oracle")
        _ORACLE = CoreLib(ORACLE_LIB)
    return _ORACLE


# ------------------------------------------------------------ NumPy restatements

def local_index_msb(i: int, qubits: Sequence[int]) -> int:
    """Local row index of global index i: qubits[0] is the most significant bit."""
    k = len(qubits)
    return sum(((i >> q) & 1) << (k - 1 - j) for j, q in enumerate(qubits))


def full_operator(n: int, qubits: Sequence[int], U) -> np.ndarray:
    """Brute-force 2^n x 2^n operator of U on `qubits` (MSB-first local index).

    Entry (i, j) is U[loc(i), loc(j)] when i and j agree on every other qubit.
    Quadratic in 2^n; for n <= 10.
    """
    U = np.asarray(U, dtype=np.complex128)
    dim = 1 << n
    mask = sum(1 << q for q in qubits)
    idx = np.arange(dim)
    loc = np.zeros(dim, dtype=np.int64)
    for j, q in enumerate(qubits):
        loc |= ((idx >> q) & 1) << (len(qubits) - 1 - j)
    rest = idx & ~mask
    same = rest[:, None] == rest[None, :]
    return np.where(same, U[loc[:, None], loc[None, :]], 0.0)


def apply_tensor(state: np.ndarray, n: int, qubits: Sequence[int], U) -> np.ndarray:
    """Tensor-axis application, the scheme of invariantq/core/binding.py:603-620."""
    k = len(qubits)
    U = np.asarray(U, dtype=np.complex128)
    rest = [q for q in range(n) if q not in qubits]
    axes = list(qubits) + rest
    t = np.asarray(state, dtype=np.complex128).reshape([2] * n, order="F")
    moved = np.transpose(t, axes).reshape(1 << k, -1)
    out = (U @ moved).reshape([2] * n)
    return np.transpose(out, np.argsort(axes)).reshape(-1, order="F")


def haar_unitary(rng: np.random.Generator, d: int) -> np.ndarray:
    """Haar-random U(d): QR of a complex Ginibre matrix with the R-diagonal phase fix (Mezzadri 2007)."""
    z = (rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))) / math.sqrt(2.0)
    q, r = np.linalg.qr(z)
    ph = np.diagonal(r) / np.abs(np.diagonal(r))
    return q * ph[None, :]


def seed_tree(master: int, job: str, index: int) -> int:
    """Per-task seed, invariantq/distributed/planner.py:110-115."""
    digest = hashlib.sha256(f"{master}:{job}:{index}".encode("utf-8")).digest()
    return int.from_bytes(digest[:8], byteorder="big")


PAULI = {
    "I": np.eye(2, dtype=np.complex128),
    "X": np.array([[0, 1], [1, 0]], dtype=np.complex128),
    "Y": np.array([[0, -1j], [1j, 0]], dtype=np.complex128),
    "Z": np.array([[1, 0], [0, -1]], dtype=np.complex128),
}


def pauli_expval_dense(state: np.ndarray, n: int, ops: str, qubits: Sequence[int]) -> complex:
    """<psi| P |psi> by applying each Pauli factor as a dense 1q operator."""
    out = np.asarray(state, dtype=np.complex128).copy()
    for c, q in zip(ops, qubits):
        out = apply_tensor(out, n, [q], PAULI.get(c, PAULI["I"]))
    return complex(np.vdot(state, out))


def tfim_terms(n: int, hx: float = 0.5) -> List[Tuple[float, str, List[int]]]:
    """H = sum_i Z_i Z_{i+1} + hx * sum_i X_i on an open chain."""
    terms = [(1.0, "ZZ", [i, i + 1]) for i in range(n - 1)]
    terms += [(hx, "X", [i]) for i in range(n)]
    return terms


def tfim_diag_and_x(n: int, hx: float = 0.5) -> Tuple[np.ndarray, float]:
    idx = np.arange(1 << n)
    z = 1 - 2 * ((idx[:, None] >> np.arange(n)[None, :]) & 1)
    diag = (z[:, :-1] * z[:, 1:]).sum(axis=1).astype(np.float64) if n > 1 else np.zeros(1 << n)
    return diag, hx


def tfim_apply(state: np.ndarray, n: int, hx: float = 0.5) -> np.ndarray:
    """H|psi> for the open-chain TFIM, by index arithmetic."""
    diag, _ = tfim_diag_and_x(n, hx)
    idx = np.arange(1 << n)
    out = diag * state
    for q in range(n):
        out = out + hx * state[idx ^ (1 << q)]
    return out


def tfim_dense(n: int, hx: float = 0.5) -> np.ndarray:
    """Dense H from Kronecker products (qubit q is bit q of the index)."""
    def kron_on(ops_by_qubit):
        m = np.array([[1.0 + 0j]])
        for q in reversed(range(n)):
            m = np.kron(m, ops_by_qubit.get(q, PAULI["I"]))
        return m
    H = np.zeros((1 << n, 1 << n), dtype=np.complex128)
    for i in range(n - 1):
        H += kron_on({i: PAULI["Z"], i + 1: PAULI["Z"]})
    for i in range(n):
        H += hx * kron_on({i: PAULI["X"]})
    return H


# ------------------------------------------------ parametric circuits (gradients)

def _ry(t):
    c, s = math.cos(t / 2), math.sin(t / 2)
    return np.array([[c, -s], [s, c]], dtype=np.complex128)


def _rz(t):
    return np.array([[np.exp(-0.5j * t), 0], [0, np.exp(0.5j * t)]], dtype=np.complex128)


def _rx(t):
    c, s = math.cos(t / 2), math.sin(t / 2)
    return np.array([[c, -1j * s], [-1j * s, c]], dtype=np.complex128)


_CZ = np.diag([1, 1, 1, -1]).astype(np.complex128)


def hea_ops(n: int, layers: int, params, ring: bool = True):
    """RY on every qubit, RZ on every qubit, then CZ on (q, q+1) and (n-1, 0), per layer.
    Parameter layout: layer l uses params[2nl + q] (RY) and params[2nl + n + q] (RZ)."""
    p = np.asarray(params, dtype=np.float64)
    ops = []
    edges = [(q, q + 1) for q in range(n - 1)] + ([(n - 1, 0)] if ring and n > 2 else [])
    for l in range(layers):
        b = 2 * n * l
        ops += [([q], _ry(p[b + q])) for q in range(n)]
        ops += [([q], _rz(p[b + n + q])) for q in range(n)]
        ops += [([a, c], _CZ) for a, c in edges]
    return ops


def energy(core: CoreLib, n: int, ops, terms) -> float:
    """sum_t c_t Re <psi|P_t|psi> for psi = ops|0>, on the oracle (complex128)."""
    sv = core.statevector(n)
    for qs, m in ops:
        (sv.apply_1q(qs[0], m) if len(qs) == 1 else sv.apply_2q(qs[0], qs[1], m))
    return float(sum(c * sv.expectation_value_pauli(o, q).real for c, o, q in terms))


def grad_parameter_shift(core: CoreLib, n: int, make_ops, terms, params, which=None) -> np.ndarray:
    """FP64 parameter-shift rule, exact for one Pauli rotation per parameter."""
    p = np.asarray(params, dtype=np.float64)
    idx = range(p.size) if which is None else which
    out = []
    for j in idx:
        a, b = p.copy(), p.copy()
        a[j] += math.pi / 2
        b[j] -= math.pi / 2
        out.append(0.5 * (energy(core, n, make_ops(a), terms) - energy(core, n, make_ops(b), terms)))
    return np.asarray(out)


def grad_central_difference(core: CoreLib, n: int, make_ops, terms, params, h: float = 1e-6, which=None) -> np.ndarray:
    p = np.asarray(params, dtype=np.float64)
    idx = range(p.size) if which is None else which
    out = []
    for j in idx:
        a, b = p.copy(), p.copy()
        a[j] += h
        b[j] -= h
        out.append((energy(core, n, make_ops(a), terms) - energy(core, n, make_ops(b), terms)) / (2 * h))
    return np.asarray(out)
