"""Parametric circuits on the GPU with adjoint-method gradients.

``GpuCircuit`` holds a gate tape and a Pauli-sum observable inside the
libinvariantq.so circuit handle (iq_circuit_*); ``AdjointExpectation`` is the
torch.autograd.Function around it: forward returns <H>(params), backward
multiplies the adjoint gradient by the incoming gradient.  Parameters may live
on the CPU or on the GPU; they are read as float64 on the host (the tape's
2x2 matrices are built there) and the result lands on the parameters' device.

There is no reference counterpart (InvariantQ optimises its VQE gradient-free,
invariantq/benchmarks/vqe.py:81-98).  Gate definitions are those of
invariantq/ir/qir.py:90-136; the Pauli convention is that of
core/src/statevector.cpp:258-291.
"""

from __future__ import annotations

import ctypes
from typing import Iterable, List, Optional, Sequence, Tuple

import numpy as np

from .core import binding
from .workloads import Ansatz

RX, RY, RZ, U3, CZ, CX, MATRIX1, MATRIX2 = range(8)
_H_MATRIX = np.array([[1, 1], [1, -1]], dtype=np.complex128) / np.sqrt(2.0)
_CX_MATRIX = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], dtype=np.complex128)
_CZ_MATRIX = np.diag([1, 1, 1, -1]).astype(np.complex128)
_KIND = {"rx": RX, "ry": RY, "rz": RZ, "u3": U3, "cz": CZ, "cx": CX, "cnot": CX}

_d = ctypes.POINTER(ctypes.c_double)
_s = ctypes.POINTER(ctypes.c_size_t)
_i64 = ctypes.POINTER(ctypes.c_int64)


def _dp(a):
    return a.ctypes.data_as(_d)


class GpuCircuit:
    """A gate tape + observable evaluated on the GPU (complex128 or complex64 state)."""

    def __init__(self, num_qubits: int, precision: str = "complex64"):
        L = binding.lib()
        self.num_qubits = num_qubits
        self.precision = precision
        self.num_params = 0
        L.iq_clear_error()
        self._free = L.iq_circuit_free  # held here: module globals may be gone when __del__ runs at exit
        self._ptr = L.iq_circuit_create(num_qubits, binding.PRECISIONS[precision])
        if not self._ptr:
            binding.check_error()
            raise MemoryError("could not allocate the circuit states on the GPU")

    def __del__(self):
        if getattr(self, "_ptr", None):
            self._free(self._ptr)
            self._ptr = None

    # ---------------------------------------------------------------- tape
    def add(self, kind, qubits: Sequence[int], params: Sequence[int] = (), matrix=None) -> int:
        if isinstance(kind, str):
            kind = _KIND[kind.lower()]
        q0 = int(qubits[0])
        q1 = int(qubits[1]) if len(qubits) > 1 else 0
        pidx = np.full(3, -1, dtype=np.int64)
        pidx[: len(params)] = params
        for p in params:
            self.num_params = max(self.num_params, int(p) + 1)
        re = im = None
        if matrix is not None:
            m = np.asarray(matrix, dtype=np.complex128).reshape(-1)
            re = np.ascontiguousarray(m.real)
            im = np.ascontiguousarray(m.imag)
        r = binding._NATIVE_LIB.iq_circuit_add_gate(self._ptr, kind, q0, q1, pidx.ctypes.data_as(_i64),
                                                    _dp(re) if re is not None else None,
                                                    _dp(im) if im is not None else None)
        binding.check_error()
        return r

    @classmethod
    def from_gates(cls, num_qubits: int, gate_list, precision: str = "complex128") -> "GpuCircuit":
        """A fixed gate list [(qubits, matrix)] (1- and 2-qubit unitaries, the
        reference's local index: qubits[0] is the high bit) as a tape for the
        tiled sweep.  One-qubit gates keep their exact matrices; CX becomes
        H CZ H on its target (CZ is a sign the tile passes apply with the
        runs around it); CZ stays CZ; other two-qubit gates are fixed 4x4
        matrices, one pass each."""
        c = cls(num_qubits, precision)
        for qubits, m in gate_list:
            m = np.asarray(m, dtype=np.complex128)
            if len(qubits) == 1:
                c.add(MATRIX1, qubits, matrix=m)
            elif len(qubits) == 2:
                a, b = int(qubits[0]), int(qubits[1])
                if np.array_equal(m, _CX_MATRIX):
                    c.add(MATRIX1, [b], matrix=_H_MATRIX)
                    c.add(CZ, [a, b])
                    c.add(MATRIX1, [b], matrix=_H_MATRIX)
                elif np.array_equal(m, _CZ_MATRIX):
                    c.add(CZ, [a, b])
                else:
                    c.add(MATRIX2, [a, b], matrix=m)
            else:
                raise ValueError("the tiled sweep takes one- and two-qubit gates")
        return c

    def run(self, params=None) -> None:
        """The forward sweep alone: the circuit's state becomes U(params)|0..0>."""
        p = self._params(np.zeros(0) if params is None else params)
        binding._NATIVE_LIB.iq_circuit_run(self._ptr, _dp(p) if p.size else None, p.size)
        binding.check_error()

    def copy_state_to(self, sv) -> None:
        """Copies the circuit's state into a NativeStateVector of the same size and precision."""
        binding._NATIVE_LIB.iq_circuit_copy_state(self._ptr, sv._ptr)
        binding.check_error()

    @classmethod
    def from_ansatz(cls, ansatz: Ansatz, precision: str = "complex64") -> "GpuCircuit":
        """RY then RZ on every qubit and a CZ ring per layer (workloads.Ansatz parameter layout)."""
        c = cls(ansatz.n, precision)
        n = ansatz.n
        for l in range(ansatz.layers):
            base = 2 * n * l
            for q in range(n):
                c.add(RY, [q], [base + q])
            for q in range(n):
                c.add(RZ, [q], [base + n + q])
            for a, b in ansatz.edges():
                c.add(CZ, [a, b])
        c.num_params = ansatz.num_params
        return c

    def set_observable(self, terms: Iterable[Tuple[float, str, Sequence[int]]]) -> None:
        terms = list(terms)
        coeffs = np.ascontiguousarray([float(c) for c, _, _ in terms])
        ops = "".join(o for _, o, _ in terms).encode("ascii")
        qubits = np.ascontiguousarray([q for _, _, qs in terms for q in qs], dtype=np.uint64)
        lens = np.ascontiguousarray([len(o) for _, o, _ in terms], dtype=np.uint64)
        binding._NATIVE_LIB.iq_circuit_set_observable(self._ptr, len(terms), _dp(coeffs), ops, qubits.ctypes.data_as(_s),
                                                      lens.ctypes.data_as(_s))
        binding.check_error()

    @property
    def num_gates(self) -> int:
        return binding._NATIVE_LIB.iq_circuit_num_gates(self._ptr)

    @property
    def num_segments(self) -> int:
        return binding._NATIVE_LIB.iq_circuit_num_segments(self._ptr)

    @property
    def stream(self) -> int:
        return binding._NATIVE_LIB.iq_circuit_stream(self._ptr) or 0

    def set_stored_budget(self, nbytes: int) -> None:
        """Device memory a gradient may keep for the forward sweep's pass
        outputs (default 16 GiB); within it the reverse sweep reads psi back
        instead of recomputing it.  0 turns the stored states off."""
        if isinstance(nbytes, bool) or not isinstance(nbytes, int) or nbytes < 0:
            raise ValueError("nbytes must be a non-negative integer")
        if binding._NATIVE_LIB.iq_circuit_set_stored_budget(self._ptr, nbytes) != 0:
            binding.check_error()

    def set_lambda_tiles(self, on: bool) -> None:
        """Reverse passes over stored outputs: lambda alone in LDS with psi read
        from its slot (True, the default), or both staged (False)."""
        binding._NATIVE_LIB.iq_circuit_set_lambda_tiles(self._ptr, 1 if on else 0)

    LAUNCH_KINDS = ("forward", "observable", "reverse", "two_qubit", "finalize")

    def set_timing(self, on: bool) -> None:
        """Bracket every launch of the tiled sweeps (16 qubits and up) with
        HIP events on the circuit's stream (a profiling aid; off by default)."""
        binding._NATIVE_LIB.iq_circuit_set_timing(self._ptr, 1 if on else 0)

    def launch_times(self):
        """[(kind, ms)] of the last expectation or gradient's launches, in
        launch order, while timing is on."""
        n = binding._NATIVE_LIB.iq_circuit_launch_times(self._ptr, None, None, 0)
        kinds = (ctypes.c_int * max(n, 1))()
        ms = np.empty(max(n, 1), dtype=np.float64)
        n = binding._NATIVE_LIB.iq_circuit_launch_times(self._ptr, kinds, _dp(ms), n)
        return [(self.LAUNCH_KINDS[kinds[i]], float(ms[i])) for i in range(n)]

    # ------------------------------------------------------------ evaluate
    def _params(self, params) -> np.ndarray:
        p = np.ascontiguousarray(np.asarray(params, dtype=np.float64).reshape(-1))
        if p.size < self.num_params:
            raise ValueError(f"circuit uses {self.num_params} parameters, got {p.size}")
        return p

    def expectation(self, params) -> float:
        p = self._params(params)
        e = binding._NATIVE_LIB.iq_circuit_expectation(self._ptr, _dp(p), p.size)
        binding.check_error()
        return e

    def gradient(self, params) -> Tuple[float, np.ndarray]:
        """Adjoint method: (<H>, d<H>/dparams)."""
        p = self._params(params)
        g = np.empty(p.size, dtype=np.float64)
        e = ctypes.c_double(0.0)
        binding._NATIVE_LIB.iq_circuit_gradient(self._ptr, _dp(p), p.size, ctypes.byref(e), _dp(g))
        binding.check_error()
        return e.value, g

    def parameter_shift(self, params, which: Optional[Sequence[int]] = None) -> np.ndarray:
        """Parameter-shift rule on the same engine for the listed parameters (all by default)."""
        p = self._params(params)
        if which is None:
            w = None
            count = p.size
        else:
            w = np.ascontiguousarray(which, dtype=np.int64)
            count = w.size
        g = np.empty(max(count, 1), dtype=np.float64)
        binding._NATIVE_LIB.iq_circuit_parameter_shift(self._ptr, _dp(p), p.size,
                                                       w.ctypes.data_as(_i64) if w is not None else None,
                                                       0 if w is None else w.size, _dp(g))
        binding.check_error()
        return g[:count]

    def amplitudes(self) -> np.ndarray:
        """The state after the last evaluation: the circuit's output after
        `expectation` or `parameter_shift`; after `gradient`, the output when
        the pass outputs were stored, else |0..0> up to rounding."""
        n = 1 << self.num_qubits
        re, im = np.empty(n), np.empty(n)
        binding._NATIVE_LIB.iq_circuit_get_amplitudes(self._ptr, _dp(re), _dp(im))
        binding.check_error()
        return re + 1j * im


def _autograd_function():
    import torch

    class AdjointExpectation(torch.autograd.Function):
        """E = <0|U(theta)^dag H U(theta)|0>; backward uses the adjoint gradient."""

        @staticmethod
        def forward(ctx, params, circuit: GpuCircuit):
            host = params.detach().to("cpu", torch.float64).numpy()
            if params.requires_grad:
                e, g = circuit.gradient(host)
                ctx.save_for_backward(torch.from_numpy(g).to(params.device, params.dtype))
            else:
                e = circuit.expectation(host)
                ctx.save_for_backward(torch.zeros_like(params))
            return torch.tensor(e, dtype=params.dtype, device=params.device)

        @staticmethod
        def backward(ctx, grad_out):
            (g,) = ctx.saved_tensors
            return grad_out * g, None

    return AdjointExpectation


_FN = None


def expectation(circuit: GpuCircuit, params):
    """Differentiable <H>(params) for a torch tensor of parameters."""
    global _FN
    if _FN is None:
        _FN = _autograd_function()
    return _FN.apply(params, circuit)
