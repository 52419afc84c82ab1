"""Gate and channel fusion: many small operations into few wide ones.

Every operation of a statevector or density-matrix circuit costs one full
pass over the state in HBM whatever its width, so the pass count, not the
arithmetic, sets the run time.  This module rewrites an operation list into
fewer, wider operations with the same effect:

* statevector: gates on up to ``width`` (default 5) qubits, each applied with
  ``apply_fused``; five-qubit blocks run on the matrix-core kernel.  This is
  the fusion producer the reference lacks: its ``GateFusionPass`` stops at two
  qubits and lifts wider unitaries to the identity
  (invariantq/compiler/gate_fusion_pass.py:13,136-154);
* density matrix: unitaries, Kraus channels and resets on up to ``width``
  (default 2) qubits, composed as Liouville superoperators and applied with
  ``iq_density_matrix_apply_superop``.  The reference applies each gate and
  each channel as its own full pass (invariantq/simulator/density_matrix.py:35-58).

Scheduler: open buckets on pairwise disjoint qubit sets.  Operations on
disjoint qubits commute, so an operation is composed into the buckets it
touches when their union fits the width.  Otherwise it keeps those of them
whose union with it is widest within the width and the others are emitted
first (emitting all of them, the fused 20-qubit, 20-layer U3 + brick CX
circuit took 89 five-qubit passes; this way, 81).  An operation touching no
kept bucket opens its own, and the buckets still open at the end are packed
together, widest first, as far as the width allows (joining the fullest
bucket at once instead: 81 passes on that circuit; this way, 46).  Order is
preserved on every qubit.

Conventions (those of the engine): matrices are row-major, local index
MSB-first (the first listed wire is the high bit).  A density-matrix
superoperator on qubits (q0..q_{k-1}) acts on the local index (r_q0 .. r_q{k-1},
c_q0 .. c_q{k-1}) with S = sum_K K (x) conj(K) (rho -> K rho K^dagger), the
layout of the engine's own superoperator passes.
"""

from __future__ import annotations

import functools
import itertools
from dataclasses import dataclass, field
from typing import Callable, Dict, Hashable, List, Optional, Sequence, Tuple

import numpy as np

from . import gates
from .workloads import Op

Wire = Hashable


def embed(m: np.ndarray, wires: Sequence[Wire], target: Sequence[Wire]) -> np.ndarray:
    """Matrix on `wires` (MSB-first) as a matrix on `target` (a superset, MSB-first),
    the identity on the extra wires."""
    wires, target = list(wires), list(target)
    if wires == target:
        return m
    k, kt = len(wires), len(target)
    extra = [w for w in target if w not in wires]
    full = np.kron(m, np.eye(1 << (kt - k), dtype=m.dtype))  # wire order: wires + extra
    order = wires + extra
    perm = [order.index(w) for w in target]
    t = full.reshape((2,) * (2 * kt))
    t = t.transpose(perm + [kt + p for p in perm])
    return t.reshape(1 << kt, 1 << kt)


@functools.lru_cache(maxsize=4096)
def _sub_index(kt: int, pos: Tuple[int, ...]):
    """Index maps of a sub-operator on bit positions `pos` (its wires in
    MSB-first order) of a kt-bit local index i: sub(i), the sub-operator's own
    index of i; rest(i), i with those bits cleared; dep(j), the sub-index j
    deposited onto those bits.  Read-only arrays, shared by every caller."""
    i = np.arange(1 << kt, dtype=np.intp)
    k = len(pos)
    j = np.arange(1 << k, dtype=np.intp)
    sub = np.zeros(1 << kt, dtype=np.intp)
    dep = np.zeros(1 << k, dtype=np.intp)
    mask = 0
    for a, p in enumerate(pos):
        sub |= ((i >> p) & 1) << (k - 1 - a)
        dep |= ((j >> (k - 1 - a)) & 1) << p
        mask |= 1 << p
    rest = i & ~mask
    rows = rest[:, None] | dep[None, :]  # row i of (m on pos) @ A mixes rows rows[i, :] of A
    same = rest[:, None] == rest[None, :]
    for arr in (sub, rest, dep, rows, same):
        arr.setflags(write=False)
    return sub, rows, same


def _positions(wires: Sequence[Wire], target: Sequence[Wire]) -> Tuple[int, ...]:
    kt = len(target)
    where = {w: n for n, w in enumerate(target)}
    return tuple(kt - 1 - where[w] for w in wires)


def embed_fast(m: np.ndarray, wires: Sequence[Wire], target: Sequence[Wire]) -> np.ndarray:
    """`embed` through index maps: E[i, j] = m[sub(i), sub(j)] where i and j
    agree off the wires, else 0."""
    wires, target = list(wires), list(target)
    if wires == target:
        return m
    sub, _, same = _sub_index(len(target), _positions(wires, target))
    return m[sub[:, None], sub[None, :]] * same


def left_apply(m: np.ndarray, wires: Sequence[Wire], a: np.ndarray, target: Sequence[Wire]) -> np.ndarray:
    """embed(m, wires, target) @ a without forming the embedded operator:
    row i is sum_j m[sub(i), j] a[rest(i) | dep(j)]."""
    wires, target = list(wires), list(target)
    if wires == target:
        return m @ a
    sub, rows, _ = _sub_index(len(target), _positions(wires, target))
    return np.matmul(m[sub][:, None, :], a[rows])[:, 0, :]


def superop(kraus: Sequence[np.ndarray]) -> np.ndarray:
    """sum_K K (x) conj(K): rho -> sum K rho K^dagger on the (row bits, column bits) index."""
    ks = [np.asarray(k, dtype=np.complex128) for k in kraus]
    return sum(np.kron(k, k.conj()) for k in ks)


def dm_wires(qubits: Sequence[int]) -> List[Wire]:
    """Row wires then column wires of density-matrix qubits."""
    return [("r", q) for q in qubits] + [("c", q) for q in qubits]


RESET_KRAUS = [np.array([[1, 0], [0, 0]], dtype=np.complex128), np.array([[0, 1], [0, 0]], dtype=np.complex128)]


@dataclass
class Fused:
    """One output operation: `matrix` on `qubits` (MSB-first); `count` source ops."""
    qubits: Tuple[int, ...]
    matrix: np.ndarray
    count: int = 1


@dataclass
class _Bucket:
    qubits: List[int]
    matrix: np.ndarray
    count: int
    wires: Callable[[Sequence[int]], List[Wire]] = field(repr=False)

    def compose(self, qubits: List[int], m: np.ndarray, wires_of, count: int = 1) -> None:
        """Apply (qubits, m) after this bucket's content."""
        union = self.qubits + [q for q in qubits if q not in self.qubits]
        uw = wires_of(union)
        a = embed_fast(self.matrix, wires_of(self.qubits), uw)
        self.matrix = left_apply(m, wires_of(qubits), a, uw)
        self.qubits = union
        self.count += count


class Fuser:
    """Bucket scheduler over (qubits, matrix) operations of one kind."""

    def __init__(self, width: int, wires_of: Callable[[Sequence[int]], List[Wire]]):
        if width < 1:
            raise ValueError("fusion width must be at least 1")
        self.width = width
        self.wires_of = wires_of
        self.open: List[_Bucket] = []
        self.out: List[Fused] = []

    def _emit(self, b: _Bucket) -> None:
        self.open.remove(b)
        self.out.append(Fused(tuple(b.qubits), b.matrix, b.count))

    def add(self, qubits: Sequence[int], m: np.ndarray) -> None:
        qs = list(qubits)
        if len(set(qs)) != len(qs):
            raise ValueError(f"repeated qubit in {qs}")
        if len(qs) > self.width:
            raise ValueError(f"operation on {len(qs)} qubits exceeds the fusion width {self.width}")
        sq = set(qs)
        hit = [b for b in self.open if sq & set(b.qubits)]
        # keep the touched buckets whose union with the operation is widest
        # within the width (all of them when it fits), emit the others first
        keep, widest = (), 0
        for r in range(len(hit), 0, -1):
            for sel in itertools.combinations(hit, r):
                u = len(sq.union(*(b.qubits for b in sel)))
                if widest < u <= self.width:
                    keep, widest = sel, u
            if keep:
                break
        for b in hit:
            if b not in keep:
                self._emit(b)
        if keep:
            first = keep[0]
            for b in keep[1:]:  # disjoint buckets commute: merge them into the first
                first.compose(b.qubits, b.matrix, self.wires_of, b.count)
                self.open.remove(b)
            first.compose(qs, m, self.wires_of)
            return
        # its own bucket: joining an unrelated one now would fill it before
        # the gates that connect its qubits arrive; unrelated buckets are
        # packed together when they are flushed
        self.open.append(_Bucket(list(qs), np.asarray(m, dtype=np.complex128), 1, self.wires_of))

    def flush(self) -> List[Fused]:
        # pack the open buckets (pairwise disjoint, so they commute) into as
        # few as fit the width: first fit, widest first
        packed: List[_Bucket] = []
        for b in sorted(self.open, key=lambda b: -len(b.qubits)):
            for t in packed:
                if len(t.qubits) + len(b.qubits) <= self.width:
                    t.compose(b.qubits, b.matrix, self.wires_of, b.count)
                    break
            else:
                packed.append(b)
        self.open = packed
        for b in list(self.open):
            self._emit(b)
        out, self.out = self.out, []
        return out


def fuse_statevector(ops: Sequence, width: int = 5) -> List[Fused]:
    """Gates (``workloads.Op`` or ``("gate", name, qubits, matrix)``) -> fused gates on
    at most `width` qubits.  Resets and channels are not unitary gates: fuse the
    stretches between them separately."""
    f = Fuser(width, lambda qs: list(qs))
    for op in ops:
        if isinstance(op, Op):
            if op.kind != "gate":
                raise ValueError("statevector fusion takes unitary gates only")
            qs, m = op.qubits, op.matrix if op.matrix is not None else gates.gate_matrix(op.name, op.params)
        else:
            kind, name, qs, m = op
            if kind != "gate":
                raise ValueError("statevector fusion takes unitary gates only")
            if m is None:
                m = gates.gate_matrix(name)
        f.add(list(qs), np.asarray(m, dtype=np.complex128))
    return f.flush()


def _kraus_items(qs: List[int], ks) -> List[Tuple[List[int], np.ndarray]]:
    """A channel listed on `qs`: one channel on all of them when the Kraus
    operators span them, else the 1-qubit channel on each (the caller loop of
    invariantq/simulator/density_matrix.py:54-58)."""
    d = np.asarray(ks[0]).shape[0]
    if d == 1 << len(qs):
        return [(list(qs), superop(ks))]
    if d == 2:
        s1 = superop(ks)
        return [([q], s1) for q in qs]
    raise ValueError(f"Kraus operators of size {d} do not fit qubits {qs}")


def dm_superops_of(op) -> List[Tuple[List[int], np.ndarray]]:
    """(qubits, superoperator) items of one density-matrix operation: a gate
    (unitary), a Kraus channel, a reset (K0 = |0><0|, K1 = |0><1|,
    density_matrix.cpp:258-270)."""
    if isinstance(op, Op):
        if op.kind == "kraus":
            return _kraus_items(list(op.qubits), op.kraus)
        m = op.matrix if op.matrix is not None else gates.gate_matrix(op.name, op.params)
        return [(list(op.qubits), superop([m]))]
    kind = op[0]
    if kind == "gate":
        _, name, qs, m = op
        return [(list(qs), superop([m if m is not None else gates.gate_matrix(name)]))]
    if kind == "reset":
        return [([op[1]], superop(RESET_KRAUS))]
    if kind == "noise":
        _, channel, qs, params = op
        return _kraus_items(list(qs), gates.build_kraus_channel(channel, params))
    raise ValueError(f"unknown op {op!r}")


def fuse_density_matrix(ops: Sequence, width: int = 2,
                        noise_model: Optional[Dict[str, List[np.ndarray]]] = None) -> List[Fused]:
    """Density-matrix operations -> superoperators on at most `width` qubits.

    With `noise_model`, each named gate is followed by its channel the way
    ``DensityMatrixSimulator.run`` applies it (density_matrix.py:36-50): on the
    target of a 1-qubit gate; on both targets of a 2-qubit gate, as one 2-qubit
    channel when the Kraus operators are 4x4, else the 1-qubit channel on each."""
    f = Fuser(width, dm_wires)
    for op in ops:
        items = dm_superops_of(op)
        for qs, s in items:
            f.add(qs, s)
        is_gate = (op.kind == "gate") if isinstance(op, Op) else op[0] == "gate"
        name = (op.name if isinstance(op, Op) else op[1]) if is_gate else None
        noise = noise_model.get(name) if (noise_model and is_gate) else None
        qs = items[0][0]
        if noise:
            if len(qs) == 2 and np.asarray(noise[0]).shape == (4, 4):
                f.add(qs, superop(noise))
            else:
                for q in qs:
                    f.add([q], superop(noise))
    return f.flush()
