"""Simulator callers of the engine.

Restates the op loop, validation and output formatting of
invariantq/simulator/statevector.py:21-144 (StateVectorSimulator.run,
NoisyStatevectorSimulator) and invariantq/simulator/density_matrix.py:22-112
(DensityMatrixSimulator.run) over a plain op list, so the GPU engine can be
driven the way the reference drives its native core:

* unitarity check allclose(U^dag U, I, 1e-10) unless non-unitary gates are allowed;
* dispatch by arity (1 -> apply_1q, 2 -> apply_2q, else apply_fused);
* counts keyed by MSB-first bit strings f"{index:0{n}b}"; probabilities > 1e-7 kept;
* samples from the engine's sampler (mt19937_64 per seed), or NumPy's
  default_rng(seed).choice for non-unitary runs and density matrices, as the reference does.

Optional fusion (off by default, so results follow the reference op by op):
``StateVectorSimulator(fusion_width=5)`` fuses the gates between resets into
blocks of up to 5 qubits; ``DensityMatrixSimulator(fuse=True)`` composes
gates, their noise and channels into superoperators on up to 2 qubits
(``fusion.py``).  Each fused op is one pass over the state instead of many;
results agree with the unfused run to rounding (about 1e-15 relative).
``StateVectorSimulator(tiled=True)`` runs a reset-free gate list of one- and
two-qubit gates on 16 qubits or more through the circuit engine's tile passes
(``adjoint.GpuCircuit.from_gates``): every one-qubit gate on the qubits of a
pass and the CZ signs between them in one pass over the state, CX as H CZ H,
about two passes per layer of a brick circuit; the state is then copied into
the simulator's statevector handle.  Results agree with the op-by-op run to
rounding; lists with resets take the op-by-op path.

Op forms accepted: ``workloads.Op`` objects, or tuples
``("gate", name, qubits, matrix)``, ``("reset", qubit)``,
``("noise", channel, qubits, params)``.  A gate's matrix may be None when its
name is a known gate (``gates.gate_matrix``).
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Any, Dict, List, Optional, Sequence

import numpy as np

from . import fusion, gates
from .core.binding import HAS_NATIVE_CORE, NativeDensityMatrix, NativeStateVector, native_core_status
from .workloads import Op


@dataclass
class ExecutionResult:
    """Field names of invariantq/simulator/base.py:11-22."""
    counts: Dict[str, int] = field(default_factory=dict)
    probabilities: Dict[str, float] = field(default_factory=dict)
    statevector: Optional[np.ndarray] = None
    density_matrix: Optional[np.ndarray] = None
    entanglement_entropy: Optional[float] = None
    execution_time_ms: float = 0.0
    backend_used: str = ""
    memory_used_bytes: int = 0
    shots: int = 1024
    metadata: Dict[str, Any] = field(default_factory=dict)


def checked_power_of_two(qubits: int, bytes_per_amplitude: int, *, copies: int = 1) -> int:
    """invariantq/simulator/base.py:44-53"""
    if isinstance(qubits, bool) or not isinstance(qubits, int) or qubits < 0:
        raise ValueError("qubits must be a non-negative integer")
    if copies <= 0 or bytes_per_amplitude <= 0:
        raise ValueError("allocation estimate parameters must be positive")
    if qubits >= 62:
        raise MemoryError("requested dense allocation exceeds supported integer range")
    return (1 << qubits) * bytes_per_amplitude * copies


def _normalise(op):
    if isinstance(op, Op):
        if op.kind == "kraus":
            return ("kraus", op.name, list(op.qubits), op.kraus)
        m = op.matrix if op.matrix is not None else gates.gate_matrix(op.name, op.params)
        return ("gate", op.name, list(op.qubits), m)
    kind = op[0]
    if kind == "gate":
        _, name, qubits, matrix = op
        if matrix is None:
            matrix = gates.gate_matrix(name)
        return ("gate", name, list(qubits), matrix)
    if kind == "reset":
        return ("reset", None, [op[1]], None)
    if kind == "noise":
        _, channel, qubits, params = op
        return ("kraus", channel, list(qubits), gates.build_kraus_channel(channel, params))
    raise ValueError(f"unknown op {op!r}")


def _counts(samples, n):
    """Bitstring counts keyed in order of first appearance, as the per-shot
    loop of invariantq/simulator/statevector.py:104-108 builds them (the
    distinct outcomes and their counts from np.unique, then one key each)."""
    s = np.asarray(samples, dtype=np.uint64).reshape(-1)
    if s.size == 0:
        return {}
    uniq, first, cnt = np.unique(s, return_index=True, return_counts=True)
    order = np.argsort(first, kind="stable")
    return dict(zip(_bitstrings(uniq[order], n), cnt[order].tolist()))


def _bitstrings(idx, n):
    """f"{i:0{n}b}" of every index (n <= 64), the digits made in one array:
    the big-endian bytes of each index unpacked to bits, the last n kept."""
    if idx.size == 0:
        return []
    b = np.unpackbits(np.asarray(idx, dtype=">u8").view(np.uint8).reshape(-1, 8), axis=1)[:, 64 - n:]
    raw = np.ascontiguousarray(b + np.uint8(48)).tobytes().decode("ascii")
    return [raw[i:i + n] for i in range(0, len(raw), n)]


def _probabilities(probs, n):
    """Probabilities above 1e-7 keyed by bitstring, in index order
    (statevector.py:110-113); the threshold is applied on the whole array
    first, so only the kept entries are formatted."""
    p = np.asarray(probs, dtype=np.float64).reshape(-1)
    idx = np.flatnonzero(p > 1e-7)
    return dict(zip(_bitstrings(idx, n), p[idx].tolist()))


def _unitary_flags(norm) -> List[bool]:
    """For each normalised op: a gate whose matrix is a finite unitary of its
    size, by the test of statevector.py:55-65 (np.allclose(U^H U, I,
    atol=1e-10, rtol=1e-10) on a finite 2^k x 2^k matrix), made for all gates
    of one size at once; True for the other ops."""
    flags = [True] * len(norm)
    by_d: Dict[int, List[int]] = {}
    mats: Dict[int, np.ndarray] = {}
    for i, (kind, _, qubits, matrix) in enumerate(norm):
        if kind != "gate":
            continue
        d = 1 << len(qubits)
        m = np.asarray(matrix, dtype=np.complex128)
        if m.shape != (d, d):
            flags[i] = False
            continue
        by_d.setdefault(d, []).append(i)
        mats[i] = m
    for d, idx in by_d.items():
        M = np.stack([mats[i] for i in idx])
        finite = np.isfinite(M).all(axis=(1, 2))
        with np.errstate(invalid="ignore", over="ignore"):
            P = np.conj(np.swapaxes(M, 1, 2)) @ M
            eye = np.eye(d)
            close = (np.abs(P - eye) <= 1e-10 + 1e-10 * np.abs(eye)).all(axis=(1, 2))
        for k, i in enumerate(idx):
            flags[i] = bool(finite[k] and close[k])
    return flags


class StateVectorSimulator:
    """GPU counterpart of invariantq/simulator/statevector.py:16-144."""

    def __init__(self, max_memory_bytes: Optional[int] = None, precision: str = "complex128",
                 fusion_width: Optional[int] = None, tiled: bool = False):
        self.name = "statevector"
        self.max_memory_bytes = max_memory_bytes
        self.precision = precision
        if fusion_width is not None and not 1 <= fusion_width <= 5:
            raise ValueError("fusion_width must be between 1 and 5")
        self.fusion_width = fusion_width
        self.tiled = bool(tiled)

    def evolve(self, num_qubits: int, ops: Sequence, *, allow_non_unitary: bool = False):
        """The op loop of run(): a fresh engine state with every gate and reset
        applied (fused when fusion_width is set).  Returns (state, fused op count)."""
        n = num_qubits
        sv = NativeStateVector(n, precision=self.precision)
        pending = []  # gates awaiting fusion (fusion_width set)
        fused_ops = 0

        def apply(qubits, m):
            if len(qubits) == 1:
                sv.apply_1q(qubits[0], m)
            elif len(qubits) == 2:
                sv.apply_2q(qubits[0], qubits[1], m)
            else:
                sv.apply_fused(qubits, m)

        def drain():
            nonlocal fused_ops
            if not pending:
                return
            fused = fusion.fuse_statevector(pending, self.fusion_width)
            sv.apply_fused_sequence([(f.qubits, f.matrix) for f in fused])
            fused_ops += len(fused)
            pending.clear()

        norm = [_normalise(raw) for raw in ops]
        unitary = _unitary_flags(norm)
        if (self.tiled and not allow_non_unitary and n >= 16
                and all(k == "gate" and len(q) <= 2 for k, _, q, _ in norm)):
            gl = []
            for (_, name, qubits, matrix), ok in zip(norm, unitary):
                if not ok:
                    raise ValueError(f"Gate '{name}' is not a finite unitary matrix")
                gl.append((list(qubits), np.asarray(matrix, dtype=np.complex128)))
            from .adjoint import GpuCircuit
            circ = GpuCircuit.from_gates(n, gl, self.precision)
            circ.run()
            circ.copy_state_to(sv)
            return sv, circ.num_segments
        for (kind, name, qubits, matrix), ok in zip(norm, unitary):
            if kind == "reset":
                drain()
                sv.reset_qubit(qubits[0])
                continue
            if kind != "gate":
                raise ValueError("noise channels need the density-matrix simulator")
            m = np.asarray(matrix, dtype=np.complex128)
            if not ok and not allow_non_unitary:
                raise ValueError(f"Gate '{name}' is not a finite unitary matrix")
            if self.fusion_width:
                pending.append(("gate", name, list(qubits), m))
            else:
                apply(qubits, m)
        drain()
        return sv, (fused_ops if self.fusion_width else None)

    def run(self, num_qubits: int, ops: Sequence, shots: int = 1024, seed: int = 42, *,
            allow_non_unitary: bool = False, entropy: bool = True) -> ExecutionResult:
        t0 = time.perf_counter()
        if isinstance(shots, bool) or not isinstance(shots, (int, np.integer)) or shots <= 0:
            raise ValueError("shots must be a positive integer")
        if isinstance(seed, bool) or not isinstance(seed, (int, np.integer)) or seed < 0:
            raise ValueError("seed must be a non-negative integer")
        n = num_qubits
        if n == 0:
            raise ValueError("Circuit has 0 qubits")
        mem_bytes = checked_power_of_two(n, 16, copies=2)
        if self.max_memory_bytes is not None and mem_bytes > self.max_memory_bytes:
            raise MemoryError(f"statevector requires approximately {mem_bytes} bytes including a working copy, "
                              f"above configured limit {self.max_memory_bytes}")
        sv, fused_ops = self.evolve(n, ops, allow_non_unitary=allow_non_unitary)
        amps = sv.get_amplitudes()
        probs_vec = sv.get_probabilities()
        ent = 0.0
        if entropy and n >= 2:
            ent = sv.von_neumann_entropy(list(range(n // 2))) if n // 2 <= 15 else None
        if allow_non_unitary:
            total = float(np.sum(probs_vec))
            if not np.isfinite(total) or total <= 0.0:
                raise ValueError("non-unitary trajectory produced an invalid state norm")
            samples = np.random.default_rng(seed).choice(1 << n, size=shots, p=probs_vec / total)
        else:
            samples = sv.sample(shots, seed)
        return ExecutionResult(
            counts=_counts(samples, n), probabilities=_probabilities(probs_vec, n), statevector=amps,
            entanglement_entropy=ent, execution_time_ms=(time.perf_counter() - t0) * 1e3,
            backend_used="statevector", memory_used_bytes=mem_bytes, shots=shots,
            metadata={"num_qubits": n, "dimension": 1 << n, "allow_non_unitary": allow_non_unitary,
                      "state_norm": float(np.sum(probs_vec)), "core": native_core_status(),
                      "precision": self.precision, "fusion_width": self.fusion_width, "tiled": self.tiled,
                      "fused_ops": fused_ops,
                      "provenance": {"requested_backend": self.name, "selected_backend": "statevector",
                                     "implementation": "dense_statevector", "native": HAS_NATIVE_CORE,
                                     "exact": not allow_non_unitary,
                                     "resource_limit_bytes": self.max_memory_bytes, "completion": "completed"}})


class NoisyStatevectorSimulator(StateVectorSimulator):
    """invariantq/simulator/statevector.py:147-163: non-unitary gates allowed."""

    def run(self, num_qubits: int, ops: Sequence, shots: int = 1024, seed: int = 42, **kw) -> ExecutionResult:
        kw["allow_non_unitary"] = True
        return super().run(num_qubits, ops, shots=shots, seed=seed, **kw)


class DensityMatrixSimulator:
    """GPU counterpart of invariantq/simulator/density_matrix.py:13-112."""

    def __init__(self, noise_model: Optional[Dict[str, List[np.ndarray]]] = None,
                 max_memory_bytes: Optional[int] = None, fuse: bool = False):
        self.name = "density_matrix"
        self.noise_model = noise_model
        self.max_memory_bytes = max_memory_bytes
        self.fuse = fuse

    def run(self, num_qubits: int, ops: Sequence, shots: int = 1024, seed: int = 42) -> ExecutionResult:
        t0 = time.perf_counter()
        n = num_qubits
        mem_bytes = checked_power_of_two(n, 16, copies=1 << n)
        if n > 14 or (self.max_memory_bytes is not None and mem_bytes > self.max_memory_bytes):
            raise MemoryError(f"Density matrix requires approximately {mem_bytes} bytes. "
                              f"For n > 14, please use StateVector or TensorNetwork (MPS) backend.")
        dm = NativeDensityMatrix(n)
        dim = 1 << n
        fused = None
        if self.fuse:
            for raw in ops:
                kind, name, qubits, matrix = _normalise(raw)
                if kind == "gate" and len(qubits) > 2:
                    raise ValueError("density-matrix gates act on one or two qubits")
            fused = fusion.fuse_density_matrix(ops, 2, self.noise_model)
            for f in fused:
                dm.apply_superop(f.qubits, f.matrix)
            ops = []
        for raw in ops:
            kind, name, qubits, matrix = _normalise(raw)
            if kind == "gate":
                noise = self.noise_model.get(name) if self.noise_model else None
                if len(qubits) == 1:
                    dm.apply_1q_unitary(qubits[0], matrix)
                    if noise:
                        dm.apply_1q_kraus(qubits[0], noise)
                elif len(qubits) == 2:
                    dm.apply_2q_unitary(qubits[0], qubits[1], matrix)
                    if noise:
                        if np.asarray(noise[0]).shape == (4, 4):
                            dm.apply_2q_kraus(qubits[0], qubits[1], noise)
                        else:
                            dm.apply_1q_kraus(qubits[0], noise)
                            dm.apply_1q_kraus(qubits[1], noise)
                else:
                    raise ValueError("density-matrix gates act on one or two qubits")
            elif kind == "reset":
                dm.reset_qubit(qubits[0])
            else:
                # a channel listed on two qubits whose Kraus operators are 4x4
                # (depolarizing_2q) acts on the pair; 2x2 operators act on each
                # listed qubit (density_matrix.py:54-58)
                d = np.asarray(matrix[0]).shape[0]
                if len(qubits) == 2 and d == 4:
                    dm.apply_2q_kraus(qubits[0], qubits[1], matrix)
                elif d == 2:
                    for q in qubits:
                        dm.apply_1q_kraus(q, matrix)
                else:
                    raise ValueError(f"Kraus operators of size {d} do not fit qubits {qubits}")
        mat = dm.get_matrix()
        trace, purity, ent = dm.trace(), dm.purity(), dm.von_neumann_entropy()
        diag = np.maximum(0.0, np.real(np.diag(mat)))
        s = np.sum(diag)
        if s > 0:
            diag = diag / s
        samples = np.random.default_rng(seed).choice(dim, size=shots, p=diag)
        return ExecutionResult(
            counts=_counts(samples, n), probabilities=_probabilities(diag, n), density_matrix=mat,
            entanglement_entropy=ent, execution_time_ms=(time.perf_counter() - t0) * 1e3,
            backend_used="density_matrix", memory_used_bytes=mem_bytes, shots=shots,
            metadata={"num_qubits": n, "dimension": dim, "trace": trace, "purity": purity,
                      "core": native_core_status(), "fused_ops": len(fused) if fused is not None else None,
                      "provenance": {"requested_backend": self.name, "selected_backend": "density_matrix",
                                     "implementation": "density_matrix", "native": HAS_NATIVE_CORE, "exact": True,
                                     "resource_limit_bytes": self.max_memory_bytes, "completion": "completed"}})
