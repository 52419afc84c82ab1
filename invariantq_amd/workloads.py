"""Seeded synthetic workloads for the statevector / density-matrix hot path.

Every generator uses numpy.random.default_rng(seed) (PCG64) so the same seed
gives the same circuit on every host.  Circuits are plain lists of ``Op`` so
they can drive the HIP engine and the CPU oracle through the same calls that
invariantq/simulator/statevector.py:50-79 and density_matrix.py:35-56 make.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import gates

TWO_PI = 2.0 * math.pi


@dataclass
class Op:
    kind: str                      # "gate" | "kraus"
    qubits: List[int]
    matrix: Optional[np.ndarray] = None
    kraus: Optional[List[np.ndarray]] = None
    name: str = ""
    params: Tuple[float, ...] = ()


def haar_unitary(rng: np.random.Generator, d: int) -> np.ndarray:
    """Haar-random U(d): QR of a complex Ginibre matrix with R-diagonal phases fixed (Mezzadri 2007)."""
    z = (rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))) / math.sqrt(2.0)
    q, r = np.linalg.qr(z)
    d_ = np.diagonal(r)
    return q * (d_ / np.abs(d_))[None, :]


def brick_pairs(n: int, layer: int) -> List[Tuple[int, int]]:
    """CX pairs (2i, 2i+1) on even layers, (2i+1, 2i+2) on odd layers."""
    start = layer % 2
    return [(q, q + 1) for q in range(start, n - 1, 2)]


def random_u3_cx(n: int, layers: int, seed: int) -> List[Op]:
    """U3(theta, phi, lambda ~ U[0, 2pi)) on every qubit, then brick-pattern CX, per layer."""
    rng = np.random.default_rng(seed)
    ops: List[Op] = []
    for layer in range(layers):
        for q in range(n):
            th, ph, la = (float(v) for v in rng.uniform(0.0, TWO_PI, size=3))
            ops.append(Op("gate", [q], gates.u3(th, ph, la), name="u3", params=(th, ph, la)))
        for a, b in brick_pairs(n, layer):
            ops.append(Op("gate", [a, b], gates.CX, name="cx"))
    return ops


def u3_layer(n: int, seed: int) -> List[Op]:
    """One seeded U3 layer on every qubit (every amplitude becomes non-zero)."""
    rng = np.random.default_rng(seed)
    out = []
    for q in range(n):
        th, ph, la = (float(v) for v in rng.uniform(0.0, TWO_PI, size=3))
        out.append(Op("gate", [q], gates.u3(th, ph, la), name="u3", params=(th, ph, la)))
    return out


def fused_blocks(n: int, count: int, seed: int, k: int = 5, low_span: int = 10) -> List[Op]:
    """`count` Haar-random 2^k x 2^k unitaries; even entries on a random k-subset of
    qubits [0, low_span), odd entries on a random k-subset of all n qubits."""
    rng = np.random.default_rng(seed)
    ops = []
    for i in range(count):
        span = low_span if i % 2 == 0 else n
        qs = [int(v) for v in rng.choice(span, size=k, replace=False)]
        ops.append(Op("gate", qs, haar_unitary(rng, 1 << k), name=f"fused{k}"))
    return ops


# ---------------------------------------------------- variational ansatz (HEA)

@dataclass
class Ansatz:
    """Layers of RY on every qubit, RZ on every qubit, then a CZ ring.

    Parameter j of layer l: RY angles are params[l*2n + q], RZ angles are
    params[l*2n + n + q].
    """
    n: int
    layers: int
    ring: bool = True

    @property
    def num_params(self) -> int:
        return 2 * self.n * self.layers

    def edges(self) -> List[Tuple[int, int]]:
        e = [(q, q + 1) for q in range(self.n - 1)]
        if self.ring and self.n > 2:
            e.append((self.n - 1, 0))
        return e

    def ops(self, params: Sequence[float]) -> List[Op]:
        p = np.asarray(params, dtype=np.float64)
        out: List[Op] = []
        for l in range(self.layers):
            base = 2 * self.n * l
            for q in range(self.n):
                out.append(Op("gate", [q], gates.ry(p[base + q]), name="ry", params=(float(p[base + q]),)))
            for q in range(self.n):
                t = float(p[base + self.n + q])
                out.append(Op("gate", [q], gates.rz(t), name="rz", params=(t,)))
            for a, b in self.edges():
                out.append(Op("gate", [a, b], gates.CZ, name="cz"))
        return out


def tfim_terms(n: int, hx: float = 0.5) -> List[Tuple[float, str, List[int]]]:
    """H = sum_{i<n-1} Z_i Z_{i+1} + hx * sum_i X_i (open chain)."""
    return [(1.0, "ZZ", [i, i + 1]) for i in range(n - 1)] + [(hx, "X", [i]) for i in range(n)]


def batch_specs(count: int, seed: int, nmin: int = 10, nmax: int = 20, lmin: int = 4, lmax: int = 16):
    """(n, layers, params) per circuit: n ~ U{nmin..nmax}, layers ~ U{lmin..lmax}, params ~ U[0, 2pi)."""
    rng = np.random.default_rng(seed)
    ns = rng.integers(nmin, nmax + 1, size=count)
    ls = rng.integers(lmin, lmax + 1, size=count)
    out = []
    for n, l in zip(ns, ls):
        a = Ansatz(int(n), int(l))
        out.append((int(n), int(l), rng.uniform(0.0, TWO_PI, size=a.num_params)))
    return out


# --------------------------------------------------------- noisy density matrix

def noisy_dm_circuit(n: int, depth: int, seed: int, p_depol: float = 1e-3,
                     gamma: float = 1e-3, lam: float = 5e-4) -> List[Op]:
    """Random U3 + brick-CX circuit; depolarizing on each touched qubit after every
    gate; amplitude and phase damping on every qubit at the end of each layer."""
    dep = gates.depolarizing(p_depol)
    amp = gates.amplitude_damping(gamma)
    phs = gates.phase_damping(lam)
    ops: List[Op] = []
    base = random_u3_cx(n, depth, seed)
    per_layer = []
    cur = []
    layer = 0
    count_in_layer = n + len(brick_pairs(n, 0))
    for op in base:
        cur.append(op)
        if len(cur) == count_in_layer:
            per_layer.append(cur)
            layer += 1
            cur = []
            count_in_layer = n + len(brick_pairs(n, layer))
    for layer_ops in per_layer:
        for op in layer_ops:
            ops.append(op)
            for q in op.qubits:
                ops.append(Op("kraus", [q], kraus=dep, name="depolarizing", params=(p_depol,)))
        for q in range(n):
            ops.append(Op("kraus", [q], kraus=amp, name="amplitude_damping", params=(gamma,)))
        for q in range(n):
            ops.append(Op("kraus", [q], kraus=phs, name="phase_damping", params=(lam,)))
    return ops
