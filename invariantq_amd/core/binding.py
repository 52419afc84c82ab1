"""ctypes binding of the gfx950 libinvariantq.so.

Mirrors the reference binding invariantq/core/binding.py: the same library
name and search order (:14-40), ``HAS_NATIVE_CORE`` / ``native_core_status()``
(:42-52), argtypes for every exported name (:54-299), and the
``NativeStateVector`` (:302-395) / ``NativeDensityMatrix`` (:398-476) classes
with the same methods, argument meaning and marshalling (split re/im float64,
qubits as uint64).  Differences:

* there is no NumPy fallback: without the HIP library the classes raise
  RuntimeError, as the reference does when its native core is absent;
* errors the reference raises as C++ exceptions (which abort the process)
  come back as ValueError / RuntimeError from ``iq_last_error()``;
* ``NativeStateVector(n, precision="complex64")`` selects the fast form.
"""

from __future__ import annotations

import ctypes
import os
import platform
from pathlib import Path
from typing import List, Optional, Sequence

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_DIRS = [
    _HERE,
    _HERE.parent.parent / "core",
    Path("/usr/local/lib"),
    Path("/usr/lib"),
]
_LIB_NAMES = ("libinvariantq.so",)

# One HIP runtime per process: load the HSA and HIP runtimes under their bare
# names first.  A later `import torch` (whose libraries ask for these bare
# names) then binds to the same copies instead of loading the ones it bundles;
# if torch was imported first, these calls return its already-loaded copies.
for _rt in ("libhsa-runtime64.so", "libamdhip64.so"):
    try:
        ctypes.CDLL(_rt, mode=ctypes.RTLD_GLOBAL)
    except OSError:
        pass

_NATIVE_LIB = None
_NATIVE_LOAD_ERRORS: List[str] = []
_LIB_PATH: Optional[Path] = None
for _dir in _LIB_DIRS:
    for _name in _LIB_NAMES:
        _p = _dir / _name
        if _p.exists():
            try:
                _NATIVE_LIB = ctypes.CDLL(str(_p))
                if not hasattr(_NATIVE_LIB, "iq_backend_name"):
                    _NATIVE_LOAD_ERRORS.append(f"{_p}: not the hip-gfx950 engine")
                    _NATIVE_LIB = None
                    continue
                _LIB_PATH = _p
                break
            except OSError as exc:
                _NATIVE_LOAD_ERRORS.append(f"{_p}: {exc}")
    if _NATIVE_LIB is not None:
        break

HAS_NATIVE_CORE = _NATIVE_LIB is not None

_d = ctypes.POINTER(ctypes.c_double)
_s = ctypes.POINTER(ctypes.c_size_t)
_u64 = ctypes.POINTER(ctypes.c_uint64)
_vp, _sz = ctypes.c_void_p, ctypes.c_size_t

# name -> (argtypes, restype); the first 49 names are reference c_api.h:11-65.
SIGNATURES = {
    "iq_statevector_create": ([_sz], _vp),
    "iq_statevector_free": ([_vp], None),
    "iq_statevector_reset": ([_vp], None),
    # matrix pointers of the per-gate calls as plain addresses (_vp): a
    # ctypes pointer object costs more than the launch it feeds
    "iq_statevector_apply_1q": ([_vp, _sz, _vp, _vp], None),
    "iq_statevector_apply_2q": ([_vp, _sz, _sz, _vp, _vp], None),
    "iq_statevector_apply_fused": ([_vp, _s, _sz, _d, _d], None),
    "iq_statevector_measure": ([_vp, _sz, ctypes.c_double], ctypes.c_int),
    "iq_statevector_reset_qubit": ([_vp, _sz, ctypes.c_double], None),
    "iq_statevector_measure_prob": ([_vp, _sz, ctypes.c_int], ctypes.c_double),
    "iq_statevector_get_amplitudes": ([_vp, _d, _d], None),
    "iq_statevector_get_probabilities": ([_vp, _d], None),
    "iq_statevector_sample": ([_vp, _sz, ctypes.c_uint64, _u64], None),
    "iq_statevector_entropy": ([_vp, _s, _sz], ctypes.c_double),
    "iq_statevector_expval_pauli": ([_vp, ctypes.c_char_p, _s, _sz, _d, _d], None),
    "iq_density_matrix_create": ([_sz], _vp),
    "iq_density_matrix_free": ([_vp], None),
    "iq_density_matrix_reset": ([_vp], None),
    "iq_density_matrix_apply_1q_unitary": ([_vp, _sz, _vp, _vp], None),
    "iq_density_matrix_apply_2q_unitary": ([_vp, _sz, _sz, _vp, _vp], None),
    "iq_density_matrix_apply_kraus_1q": ([_vp, _sz, _sz, _d, _d], None),
    "iq_density_matrix_get_elements": ([_vp, _d, _d], None),
    "iq_density_matrix_trace": ([_vp], ctypes.c_double),
    "iq_density_matrix_purity": ([_vp], ctypes.c_double),
    "iq_density_matrix_entropy": ([_vp], ctypes.c_double),
    "iq_density_matrix_measure": ([_vp, _sz, ctypes.c_double], ctypes.c_int),
    "iq_density_matrix_reset_qubit": ([_vp, _sz], None),
    "iq_stabilizer_create": ([_sz], _vp),
    "iq_stabilizer_free": ([_vp], None),
    "iq_stabilizer_reset": ([_vp], None),
    "iq_stabilizer_reset_qubit": ([_vp, _sz, ctypes.c_double], None),
    "iq_stabilizer_apply_h": ([_vp, _sz], None),
    "iq_stabilizer_apply_s": ([_vp, _sz], None),
    "iq_stabilizer_apply_sdg": ([_vp, _sz], None),
    "iq_stabilizer_apply_cnot": ([_vp, _sz, _sz], None),
    "iq_stabilizer_apply_cz": ([_vp, _sz, _sz], None),
    "iq_stabilizer_apply_swap": ([_vp, _sz, _sz], None),
    "iq_stabilizer_apply_x": ([_vp, _sz], None),
    "iq_stabilizer_apply_y": ([_vp, _sz], None),
    "iq_stabilizer_apply_z": ([_vp, _sz], None),
    "iq_stabilizer_measure": ([_vp, _sz, ctypes.c_double], ctypes.c_int),
    "iq_stabilizer_to_string": ([_vp, ctypes.c_char_p, _sz], None),
    "iq_mps_create": ([_sz, _sz, ctypes.c_double], _vp),
    "iq_mps_free": ([_vp], None),
    "iq_mps_reset": ([_vp], None),
    "iq_mps_apply_1q": ([_vp, _sz, _d, _d], None),
    "iq_mps_apply_2q": ([_vp, _sz, _sz, _d, _d], None),
    "iq_mps_entropy": ([_vp, _sz], ctypes.c_double),
    "iq_mps_to_statevector": ([_vp, _d, _d], None),
    "iq_mps_measure": ([_vp, _sz, ctypes.c_double], ctypes.c_int),
    # additions
    "iq_last_error": ([], ctypes.c_int),
    "iq_last_error_message": ([], ctypes.c_char_p),
    "iq_clear_error": ([], None),
    "iq_backend_name": ([], ctypes.c_char_p),
    "iq_set_gate_path": ([ctypes.c_int], None),
    "iq_set_mfma_options": ([ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int], None),
    "iq_set_entropy_options": ([ctypes.c_int], None),
    "iq_statevector_create_precision": ([_sz, ctypes.c_int], _vp),
    "iq_statevector_get_precision": ([_vp], ctypes.c_int),
    "iq_statevector_set_amplitudes": ([_vp, _d, _d], None),
    "iq_statevector_get_amplitudes_interleaved": ([_vp, _d], None),
    "iq_statevector_set_amplitudes_interleaved": ([_vp, _d], None),
    "iq_statevector_norm_squared": ([_vp], ctypes.c_double),
    "iq_statevector_normalize": ([_vp], None),
    "iq_statevector_apply_diagonal": ([_vp, _s, _sz, _d, _d], None),
    "iq_statevector_sample_uniforms": ([_vp, _sz, _d, _u64], None),
    "iq_statevector_sample_uniforms_noalloc": ([_vp, _sz, _vp, _vp], None),
    "iq_statevector_get_cdf": ([_vp, _d], None),
    "iq_statevector_expval_pauli_sum": ([_vp, _sz, _d, ctypes.c_char_p, _s, _s, _d, _d, _d], None),
    "iq_statevector_apply_fused_sequence": ([_vp, _sz, _s, _s, _d, _d], None),
    "iq_statevector_sample_tasks": ([_vp, _sz, _u64, _u64, _u64], None),
    "iq_statevector_reduced_density_matrix": ([_vp, _s, _sz, _d, _d], None),
    "iq_statevector_synchronize": ([_vp], None),
    "iq_statevector_device_ptr": ([_vp], _vp),
    "iq_statevector_stream": ([_vp], _vp),
    "iq_density_matrix_set_elements": ([_vp, _d, _d], None),
    "iq_density_matrix_get_elements_interleaved": ([_vp, _d], None),
    "iq_density_matrix_set_elements_interleaved": ([_vp, _d], None),
    "iq_density_matrix_apply_kraus": ([_vp, _s, _sz, _sz, _d, _d], None),
    "iq_density_matrix_apply_superop": ([_vp, _s, _sz, _d, _d], None),
    "iq_density_matrix_expval_z": ([_vp, _d], None),
    "iq_density_matrix_synchronize": ([_vp], None),
    "iq_density_matrix_stream": ([_vp], _vp),
    "iq_circuit_create": ([_sz, ctypes.c_int], _vp),
    "iq_circuit_free": ([_vp], None),
    "iq_circuit_add_gate": ([_vp, ctypes.c_int, _sz, _sz, ctypes.POINTER(ctypes.c_int64), _d, _d], ctypes.c_int),
    "iq_circuit_num_gates": ([_vp], _sz),
    "iq_circuit_num_segments": ([_vp], _sz),
    "iq_circuit_set_observable": ([_vp, _sz, _d, ctypes.c_char_p, _s, _s], ctypes.c_int),
    "iq_circuit_expectation": ([_vp, _d, _sz], ctypes.c_double),
    "iq_circuit_gradient": ([_vp, _d, _sz, _d, _d], ctypes.c_int),
    "iq_circuit_parameter_shift": ([_vp, _d, _sz, ctypes.POINTER(ctypes.c_int64), _sz, _d], ctypes.c_int),
    "iq_circuit_get_amplitudes": ([_vp, _d, _d], None),
    "iq_circuit_run": ([_vp, _d, _sz], ctypes.c_int),
    "iq_circuit_copy_state": ([_vp, _vp], ctypes.c_int),
    "iq_circuit_stream": ([_vp], _vp),
    "iq_circuit_set_stored_budget": ([_vp, _sz], ctypes.c_int),
    "iq_circuit_set_timing": ([_vp, ctypes.c_int], ctypes.c_int),
    "iq_circuit_set_lambda_tiles": ([_vp, ctypes.c_int], ctypes.c_int),
    "iq_circuit_launch_times": ([_vp, ctypes.POINTER(ctypes.c_int), _d, _sz], _sz),
    "iq_batch_create": ([ctypes.c_int], _vp),
    "iq_batch_free": ([_vp], None),
    "iq_batch_max_tile_bits": ([ctypes.c_int], ctypes.c_int),
    "iq_batch_max_group_width": ([ctypes.c_int], ctypes.c_int),
    "iq_batch_set_group_width": ([_vp, ctypes.c_int], ctypes.c_int),
    "iq_batch_hea_tfim": ([_vp, _sz, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32), _d, ctypes.c_double,
                           ctypes.POINTER(ctypes.c_int32), ctypes.c_int, ctypes.c_int, _d, _d, _d], ctypes.c_int),
}

REFERENCE_NAMES = list(SIGNATURES)[:49]

if HAS_NATIVE_CORE:
    for _name, (_args, _res) in SIGNATURES.items():
        _fn = getattr(_NATIVE_LIB, _name)
        _fn.argtypes = _args
        _fn.restype = _res


def native_core_status() -> dict:
    """Provenance in the shape of the reference's native_core_status() (binding.py:44-52)."""
    return {
        "active": HAS_NATIVE_CORE,
        "implementation": "hip-gfx950" if HAS_NATIVE_CORE else "unavailable",
        "platform": platform.platform(),
        "library_names": list(_LIB_NAMES),
        "library_path": str(_LIB_PATH) if _LIB_PATH else None,
        "load_errors": list(_NATIVE_LOAD_ERRORS),
    }


def lib():
    if not HAS_NATIVE_CORE:
        raise RuntimeError("the hip-gfx950 libinvariantq.so is not built (python -m invariantq_amd.build); "
                           f"load errors: {_NATIVE_LOAD_ERRORS}")
    return _NATIVE_LIB


ERR_INVALID, ERR_ZERO_PROB, ERR_OOM, ERR_DEVICE, ERR_UNSUPPORTED = 1, 2, 3, 4, 5


def check_error() -> None:
    """Raise the pending library error (if any) as a Python exception."""
    L = _NATIVE_LIB
    code = L.iq_last_error()
    if code == 0:
        return
    msg = (L.iq_last_error_message() or b"").decode("utf-8", "replace")
    L.iq_clear_error()
    if code == ERR_INVALID:
        raise ValueError(msg)
    if code == ERR_OOM:
        raise MemoryError(msg)
    raise RuntimeError(msg)


def _split(matrix, expect: Optional[int] = None):
    m = np.asarray(matrix, dtype=np.complex128).reshape(-1)
    if expect is not None and m.size != expect:
        raise ValueError(f"matrix has {m.size} entries, expected {expect}")
    return (np.ascontiguousarray(m.real, dtype=np.float64), np.ascontiguousarray(m.imag, dtype=np.float64))


def _packed(matrix, expect: int):
    """One float64 buffer: the real parts, then the imaginary parts (row-major);
    returns (buffer, address of the real parts, address of the imaginary ones)."""
    m = np.asarray(matrix, dtype=np.complex128).reshape(-1)
    if m.size != expect:
        raise ValueError(f"matrix has {m.size} entries, expected {expect}")
    buf = np.empty(2 * expect)
    buf[:expect] = m.real
    buf[expect:] = m.imag
    a = buf.ctypes.data
    return buf, a, a + 8 * expect


def _dp(a: np.ndarray):
    return a.ctypes.data_as(_d)


PRECISIONS = {"complex128": 0, "c128": 0, 0: 0, "complex64": 1, "c64": 1, 1: 1}


class NativeStateVector:
    """Device-resident statevector (binding.py:302-395 interface)."""

    def __init__(self, num_qubits: int, precision="complex128"):
        self.num_qubits = num_qubits
        self.dim = 1 << num_qubits
        L = lib()
        if precision not in PRECISIONS:
            raise ValueError(f"unknown precision {precision!r}")
        self.precision = "complex128" if PRECISIONS[precision] == 0 else "complex64"
        L.iq_clear_error()
        self._free = L.iq_statevector_free  # held here: module globals may be gone when __del__ runs at exit
        self._ptr = L.iq_statevector_create_precision(num_qubits, PRECISIONS[precision])
        if not self._ptr:
            msg = (L.iq_last_error_message() or b"").decode()
            L.iq_clear_error()
            raise MemoryError(f"Failed to allocate StateVector on the GPU: {msg}")

    def __del__(self):
        if getattr(self, "_ptr", None):
            self._free(self._ptr)
            self._ptr = None

    def reset(self):
        _NATIVE_LIB.iq_statevector_reset(self._ptr)

    def apply_1q(self, target: int, matrix):
        buf, re, im = _packed(matrix, 4)
        _NATIVE_LIB.iq_statevector_apply_1q(self._ptr, target, re, im)
        check_error()

    def apply_2q(self, q0: int, q1: int, matrix):
        buf, re, im = _packed(matrix, 16)
        _NATIVE_LIB.iq_statevector_apply_2q(self._ptr, q0, q1, re, im)
        check_error()

    def apply_fused(self, qubits: Sequence[int], matrix):
        re, im = _split(matrix, 4 ** len(qubits))
        q = np.ascontiguousarray(qubits, dtype=np.uint64)
        _NATIVE_LIB.iq_statevector_apply_fused(self._ptr, q.ctypes.data_as(_s), len(qubits), _dp(re), _dp(im))
        check_error()

    def apply_fused_sequence(self, gates_seq) -> None:
        """Apply [(qubits, matrix), ...] in one call: the gates of a fusion
        pass (each 1..5 qubits, MSB-first), matrices uploaded in one copy."""
        gl = [(list(q), np.asarray(m, dtype=np.complex128)) for q, m in gates_seq]
        if not gl:
            return
        for q, m in gl:
            if m.shape != (1 << len(q), 1 << len(q)):
                raise ValueError(f"matrix of shape {m.shape} does not fit qubits {q}")
        ks = np.ascontiguousarray([len(q) for q, _ in gl], dtype=np.uint64)
        qs = np.ascontiguousarray([x for q, _ in gl for x in q], dtype=np.uint64)
        flat = np.concatenate([m.reshape(-1) for _, m in gl])
        re, im = np.ascontiguousarray(flat.real), np.ascontiguousarray(flat.imag)
        _NATIVE_LIB.iq_statevector_apply_fused_sequence(self._ptr, len(gl), ks.ctypes.data_as(_s), qs.ctypes.data_as(_s),
                                                        _dp(re), _dp(im))
        check_error()

    def apply_diagonal(self, qubits: Sequence[int], diag):
        re, im = _split(diag, 2 ** len(qubits))
        q = np.ascontiguousarray(qubits, dtype=np.uint64)
        _NATIVE_LIB.iq_statevector_apply_diagonal(self._ptr, q.ctypes.data_as(_s), len(qubits), _dp(re), _dp(im))
        check_error()

    def measure(self, target: int, rand_val: float) -> int:
        out = _NATIVE_LIB.iq_statevector_measure(self._ptr, target, rand_val)
        check_error()
        return out

    def reset_qubit(self, target: int, rand_val: float = 0.0):
        _NATIVE_LIB.iq_statevector_reset_qubit(self._ptr, target, rand_val)
        check_error()

    def measure_prob(self, target: int, outcome: int) -> float:
        out = _NATIVE_LIB.iq_statevector_measure_prob(self._ptr, target, outcome)
        check_error()
        return out

    def get_amplitudes(self) -> np.ndarray:
        # straight into a complex128 array (the split export, c_api.h:20, then
        # re + 1j * im costs two more passes over fresh host memory)
        out = np.empty(self.dim, dtype=np.complex128)
        _NATIVE_LIB.iq_statevector_get_amplitudes_interleaved(self._ptr, _dp(out))
        check_error()
        return out

    def set_amplitudes(self, amplitudes):
        a = np.ascontiguousarray(np.asarray(amplitudes, dtype=np.complex128).reshape(-1))
        if a.size != self.dim:
            raise ValueError("amplitude count does not match the statevector dimension")
        _NATIVE_LIB.iq_statevector_set_amplitudes_interleaved(self._ptr, _dp(a))
        check_error()

    def get_probabilities(self) -> np.ndarray:
        p = np.empty(self.dim, dtype=np.float64)
        _NATIVE_LIB.iq_statevector_get_probabilities(self._ptr, _dp(p))
        check_error()
        return p

    def get_cdf(self) -> np.ndarray:
        c = np.empty(self.dim, dtype=np.float64)
        _NATIVE_LIB.iq_statevector_get_cdf(self._ptr, _dp(c))
        check_error()
        return c

    def sample(self, num_shots: int, seed: int = 42) -> np.ndarray:
        out = np.empty(num_shots, dtype=np.uint64)
        _NATIVE_LIB.iq_statevector_sample(self._ptr, num_shots, seed, out.ctypes.data_as(_u64))
        check_error()
        return out

    def sample_tasks(self, seeds, shots) -> list:
        """Shots of several tasks in one call: task t is sample(shots[t], seeds[t]);
        returns one index array per task."""
        sd = np.ascontiguousarray([int(x) for x in seeds], dtype=np.uint64)
        sh = np.ascontiguousarray([int(x) for x in shots], dtype=np.uint64)
        if sd.size != sh.size or sd.size == 0:
            raise ValueError("seeds and shots must be non-empty and of equal length")
        out = np.empty(int(sh.sum()), dtype=np.uint64)
        _NATIVE_LIB.iq_statevector_sample_tasks(self._ptr, sd.size, sd.ctypes.data_as(_u64), sh.ctypes.data_as(_u64),
                                                out.ctypes.data_as(_u64))
        check_error()
        return np.split(out, np.cumsum(sh)[:-1].astype(np.int64))

    def reduced_density_matrix(self, qubits) -> np.ndarray:
        """Partial trace onto `qubits` (bit b of a row index = qubits[b])."""
        q = np.ascontiguousarray([int(x) for x in qubits], dtype=np.uint64)
        d = 1 << q.size
        re, im = np.empty(d * d), np.empty(d * d)
        _NATIVE_LIB.iq_statevector_reduced_density_matrix(self._ptr, q.ctypes.data_as(_s), q.size, _dp(re), _dp(im))
        check_error()
        return (re + 1j * im).reshape(d, d)

    def sample_uniforms(self, uniforms) -> np.ndarray:
        u = np.ascontiguousarray(uniforms, dtype=np.float64)
        out = np.empty(u.size, dtype=np.uint64)
        _NATIVE_LIB.iq_statevector_sample_uniforms(self._ptr, u.size, _dp(u), out.ctypes.data_as(_u64))
        check_error()
        return out

    def sample_uniforms_noalloc(self, uniforms_ptr: int, results_ptr: int, num_shots: int) -> None:
        """Sampling that allocates nothing: `uniforms_ptr` (float64) and
        `results_ptr` (uint64) are device addresses of num_shots entries; the
        call only enqueues on this handle's stream (wait with synchronize())."""
        _NATIVE_LIB.iq_statevector_sample_uniforms_noalloc(self._ptr, num_shots, uniforms_ptr, results_ptr)
        check_error()

    def von_neumann_entropy(self, subsystem_qubits: Sequence[int]) -> float:
        q = np.ascontiguousarray(subsystem_qubits, dtype=np.uint64)
        out = _NATIVE_LIB.iq_statevector_entropy(self._ptr, q.ctypes.data_as(_s), len(subsystem_qubits))
        check_error()
        return out

    def expectation_value_pauli(self, pauli_str: str, qubits: Sequence[int]) -> complex:
        q = np.ascontiguousarray(qubits, dtype=np.uint64)
        vr, vi = ctypes.c_double(0.0), ctypes.c_double(0.0)
        _NATIVE_LIB.iq_statevector_expval_pauli(self._ptr, pauli_str.encode("ascii"), q.ctypes.data_as(_s), len(qubits),
                                                ctypes.byref(vr), ctypes.byref(vi))
        check_error()
        return complex(vr.value, vi.value)

    def expectation_value_pauli_sum(self, terms, per_term: bool = False):
        """terms: iterable of (coeff, ops, qubits).  Returns sum_t c_t <P_t> (and the per-term values)."""
        terms = list(terms)
        coeffs = np.ascontiguousarray([float(c) for c, _, _ in terms], dtype=np.float64)
        ops = "".join(o for _, o, _ in terms).encode("ascii")
        qubits = np.ascontiguousarray([q for _, _, qs in terms for q in qs], dtype=np.uint64)
        lens = np.ascontiguousarray([len(o) for _, o, _ in terms], dtype=np.uint64)
        vals = np.empty(2 * max(1, len(terms)), dtype=np.float64)
        vr, vi = ctypes.c_double(0.0), ctypes.c_double(0.0)
        # without per-term values the library may sum over tile passes (same
        # result to rounding)
        _NATIVE_LIB.iq_statevector_expval_pauli_sum(self._ptr, len(terms), _dp(coeffs), ops, qubits.ctypes.data_as(_s),
                                                    lens.ctypes.data_as(_s), ctypes.byref(vr), ctypes.byref(vi),
                                                    _dp(vals) if per_term else None)
        check_error()
        total = complex(vr.value, vi.value)
        if per_term:
            return total, vals[0::2][: len(terms)] + 1j * vals[1::2][: len(terms)]
        return total

    def norm_squared(self) -> float:
        out = _NATIVE_LIB.iq_statevector_norm_squared(self._ptr)
        check_error()
        return out

    def normalize(self):
        _NATIVE_LIB.iq_statevector_normalize(self._ptr)
        check_error()

    def synchronize(self):
        _NATIVE_LIB.iq_statevector_synchronize(self._ptr)
        check_error()

    @property
    def stream(self) -> int:
        return _NATIVE_LIB.iq_statevector_stream(self._ptr) or 0

    @property
    def device_ptr(self) -> int:
        return _NATIVE_LIB.iq_statevector_device_ptr(self._ptr) or 0


class NativeDensityMatrix:
    """Device-resident density matrix (binding.py:398-476 interface)."""

    def __init__(self, num_qubits: int):
        self.num_qubits = num_qubits
        self.dim = 1 << num_qubits
        L = lib()
        L.iq_clear_error()
        self._free = L.iq_density_matrix_free  # held here, as for the statevector
        self._ptr = L.iq_density_matrix_create(num_qubits)
        if not self._ptr:
            msg = (L.iq_last_error_message() or b"").decode()
            L.iq_clear_error()
            raise MemoryError(f"Failed to allocate DensityMatrix on the GPU: {msg}")

    def __del__(self):
        if getattr(self, "_ptr", None):
            self._free(self._ptr)
            self._ptr = None

    def reset(self):
        _NATIVE_LIB.iq_density_matrix_reset(self._ptr)

    def apply_1q_unitary(self, target: int, matrix):
        buf, re, im = _packed(matrix, 4)
        _NATIVE_LIB.iq_density_matrix_apply_1q_unitary(self._ptr, target, re, im)
        check_error()

    def apply_2q_unitary(self, q0: int, q1: int, matrix):
        buf, re, im = _packed(matrix, 16)
        _NATIVE_LIB.iq_density_matrix_apply_2q_unitary(self._ptr, q0, q1, re, im)
        check_error()

    def apply_1q_kraus(self, target: int, kraus_ops):
        stacked = np.stack([np.asarray(k, dtype=np.complex128).reshape(-1) for k in kraus_ops])
        re, im = _split(stacked, 4 * len(kraus_ops))
        _NATIVE_LIB.iq_density_matrix_apply_kraus_1q(self._ptr, target, len(kraus_ops), _dp(re), _dp(im))
        check_error()

    def apply_2q_kraus(self, q0: int, q1: int, kraus_ops):
        """Two-qubit Kraus channel (the reference simulator calls this name,
        invariantq/simulator/density_matrix.py:47, but its binding lacks it)."""
        stacked = np.stack([np.asarray(k, dtype=np.complex128).reshape(-1) for k in kraus_ops])
        re, im = _split(stacked, 16 * len(kraus_ops))
        q = np.ascontiguousarray([q0, q1], dtype=np.uint64)
        _NATIVE_LIB.iq_density_matrix_apply_kraus(self._ptr, q.ctypes.data_as(_s), 2, len(kraus_ops), _dp(re), _dp(im))
        check_error()

    def apply_superop(self, qubits, s):
        """Superoperator on 1 or 2 qubits: 4^k x 4^k, local index (row bits of
        `qubits`, then column bits), MSB-first; S = sum K (x) conj(K) is
        rho -> sum K rho K^dagger.  The output of ``fusion.fuse_density_matrix``."""
        qs = np.ascontiguousarray([int(q) for q in qubits], dtype=np.uint64)
        k = len(qs)
        re, im = _split(s, 1 << (4 * k))
        _NATIVE_LIB.iq_density_matrix_apply_superop(self._ptr, qs.ctypes.data_as(_s), k, _dp(re), _dp(im))
        check_error()

    def get_matrix(self) -> np.ndarray:
        out = np.empty((self.dim, self.dim), dtype=np.complex128)
        _NATIVE_LIB.iq_density_matrix_get_elements_interleaved(self._ptr, _dp(out))
        check_error()
        return out

    def set_matrix(self, rho):
        a = np.ascontiguousarray(np.asarray(rho, dtype=np.complex128).reshape(-1))
        if a.size != self.dim * self.dim:
            raise ValueError("matrix size does not match the density-matrix dimension")
        _NATIVE_LIB.iq_density_matrix_set_elements_interleaved(self._ptr, _dp(a))
        check_error()

    def trace(self) -> float:
        out = _NATIVE_LIB.iq_density_matrix_trace(self._ptr)
        check_error()
        return out

    def purity(self) -> float:
        out = _NATIVE_LIB.iq_density_matrix_purity(self._ptr)
        check_error()
        return out

    def von_neumann_entropy(self) -> float:
        out = _NATIVE_LIB.iq_density_matrix_entropy(self._ptr)
        check_error()
        return out

    def expval_z(self) -> np.ndarray:
        out = np.empty(max(1, self.num_qubits), dtype=np.float64)
        _NATIVE_LIB.iq_density_matrix_expval_z(self._ptr, _dp(out))
        check_error()
        return out[: self.num_qubits]

    def measure(self, target: int, rand_val: float) -> int:
        out = _NATIVE_LIB.iq_density_matrix_measure(self._ptr, target, rand_val)
        check_error()
        return out

    def reset_qubit(self, target: int):
        _NATIVE_LIB.iq_density_matrix_reset_qubit(self._ptr, target)
        check_error()

    def synchronize(self):
        _NATIVE_LIB.iq_density_matrix_synchronize(self._ptr)
        check_error()

    @property
    def stream(self) -> int:
        return _NATIVE_LIB.iq_density_matrix_stream(self._ptr) or 0


def set_mfma_options(batches_per_iteration: int = 0, matrix_in_lds: int = -1, nontemporal: int = -1,
                     grid_waves: int = 0, batch_rotation: int = -1) -> None:
    """Matrix-core kernel options (zeros / -1 restore the defaults)."""
    lib().iq_set_mfma_options(batches_per_iteration, matrix_in_lds, nontemporal, grid_waves, batch_rotation)


def set_entropy_options(device_jacobi_min_kept: int = 0) -> None:
    """Kept qubits from which the entropy's Jacobi runs on the device (0: default)."""
    lib().iq_set_entropy_options(device_jacobi_min_kept)


def set_gate_path(path: str) -> None:
    """Kernel selection for gate application: 'auto', 'direct', 'tiled' or 'mfma'."""
    lib().iq_set_gate_path({"auto": 0, "direct": 1, "tiled": 2, "mfma": 3}[path])
