"""Writes the golden fixtures in tests/golden/*.npz from the reference native core.

The reference core (core/src/*.cpp) is compiled from its own sources by
``make -C oracle ref`` into oracle/_ref/libinvariantq.so and driven through
ctypes with the reference binding's call shapes.  Each fixture holds an op tape
(inputs) and the reference's outputs.  Run in the build container, where
/root/reference exists:

    python tests/golden/make_golden.py

The fixtures are data only (inputs and expected outputs); nothing from the
reference's sources is stored.
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import oracle as orc  # noqa: E402
from invariantq_amd import gates, workloads  # noqa: E402
from tests.golden.replay import replay  # noqa: E402

OUT = Path(__file__).resolve().parent


class Tape:
    def __init__(self):
        self.kind, self.qubits, self.mats, self.rand = [], [], [], []

    def add(self, kind, qubits, mat=None, rand=0.0):
        self.kind.append(kind)
        self.qubits.append(list(qubits) + [-1] * (5 - len(qubits)))
        self.mats.append(np.zeros(0, np.complex128) if mat is None else np.asarray(mat, np.complex128).reshape(-1))
        self.rand.append(rand)

    def arrays(self):
        off = np.zeros(len(self.mats) + 1, dtype=np.int64)
        off[1:] = np.cumsum([m.size for m in self.mats])
        return {
            "op_kind": np.asarray(self.kind, dtype=np.int8),
            "op_qubits": np.asarray(self.qubits, dtype=np.int64),
            "op_mat_off": off,
            "mats": np.concatenate(self.mats) if self.mats else np.zeros(0, np.complex128),
            "op_rand": np.asarray(self.rand, dtype=np.float64),
        }


def circuit_tape(ops):
    t = Tape()
    for op in ops:
        t.add(0 if len(op.qubits) == 1 else 1, op.qubits, op.matrix)
    return t


PAULIS = [("ZZ", [0, 1]), ("XYZ", [2, 5, 7]), ("YY", [3, 4]), ("X", [6]), ("ZIXY", [0, 1, 2, 3]),
          ("XXXXXXXX", list(range(8))), ("YZ", [7, 0])]


def sv_u3cx_n8(ref):
    t = circuit_tape(workloads.random_u3_cx(8, 6, seed=7))
    arrays = t.arrays()
    sv = ref.statevector(8)
    replay(arrays, sv)
    seed = orc.seed_tree(42, "golden", 0)
    ev = [sv.expectation_value_pauli(p, q) for p, q in PAULIS]
    mp = [[sv.measure_prob(q, o) for o in (0, 1)] for q in range(8)]
    subsets = [[0], [0, 1], [2, 5, 7]]
    ent = [sv.von_neumann_entropy(s) for s in subsets]
    out = dict(arrays, n=8, amps=sv.get_amplitudes(), probs=sv.get_probabilities(),
               sample_seed=np.uint64(seed), samples=sv.sample(2000, seed),
               pauli_ops=np.array([p for p, _ in PAULIS]),
               pauli_qubits=np.array([q + [-1] * (8 - len(q)) for _, q in PAULIS], dtype=np.int64),
               pauli_vals=np.asarray(ev, dtype=np.complex128), measure_prob=np.asarray(mp),
               entropy_subsets=np.array([s + [-1] * (3 - len(s)) for s in subsets], dtype=np.int64),
               entropy=np.asarray(ent))
    np.savez(OUT / "sv_u3cx_n8.npz", **out)


def sv_sample_n10(ref):
    t = circuit_tape(workloads.random_u3_cx(10, 4, seed=3))
    arrays = t.arrays()
    sv = ref.statevector(10)
    replay(arrays, sv)
    seeds = [orc.seed_tree(42, "cfg0", i) for i in range(3)]
    samples = np.stack([sv.sample(10000, s) for s in seeds])
    np.savez(OUT / "sv_sample_n10.npz", **arrays, n=10, seeds=np.asarray(seeds, dtype=np.uint64),
             samples=samples, probs=sv.get_probabilities())


def sv_fused_n7(ref):
    rng = np.random.default_rng(11)
    t = Tape()
    for op in workloads.u3_layer(7, seed=5):
        t.add(0, op.qubits, op.matrix)
    for qs in ([5, 1, 3], [0, 6, 2, 4], [2, 6, 0], [6, 5, 4, 3], [4, 1], [3], [1, 0, 2, 6]):
        d = 1 << len(qs)
        t.add(2, qs, workloads.haar_unitary(rng, d))
    arrays = t.arrays()
    sv = ref.statevector(7)
    replay(arrays, sv, reverse_fused_for_lsb_first=True)
    # the same tape without reversal shows the native k=3,4 convention is LSB-first
    sv_raw = ref.statevector(7)
    replay(arrays, sv_raw, reverse_fused_for_lsb_first=False)
    np.savez(OUT / "sv_fused_n7.npz", **arrays, n=7, amps=sv.get_amplitudes(), amps_lsb_first=sv_raw.get_amplitudes())


def sv_measure_n5(ref):
    t = Tape()
    for op in workloads.random_u3_cx(5, 3, seed=9):
        t.add(0 if len(op.qubits) == 1 else 1, op.qubits, op.matrix)
    t.add(3, [2], rand=0.3)
    t.add(4, [0], rand=0.7)
    t.add(0, [1], gates.H)
    t.add(3, [4], rand=0.9)
    t.add(1, [3, 1], gates.CX)
    t.add(4, [3], rand=0.05)
    t.add(3, [1], rand=0.5)
    arrays = t.arrays()
    sv = ref.statevector(5)
    outcomes = replay(arrays, sv)
    np.savez(OUT / "sv_measure_n5.npz", **arrays, n=5, outcomes=np.asarray(outcomes), amps=sv.get_amplitudes())


def dm_tape(n, seed):
    rng = np.random.default_rng(seed)
    t = Tape()
    for q in range(n):
        th, ph, la = rng.uniform(0, 2 * np.pi, 3)
        t.add(8, [q], gates.u3(th, ph, la))
    pairs = [(0, 1), (n - 1, 1), (2, 0), (1, n - 1)]
    for a, b in pairs:
        t.add(9, [a, b], gates.CX)
        t.add(5, [a], np.stack(gates.depolarizing(0.05)))
    t.add(9, [0, 2], workloads.haar_unitary(rng, 4))
    t.add(5, [1], np.stack(gates.amplitude_damping(0.1)))
    t.add(5, [2], np.stack(gates.phase_damping(0.2)))
    t.add(5, [n - 1], np.stack(gates.bit_flip(0.1)))
    t.add(7, [1], rand=0.4)
    t.add(6, [n - 1])
    for q in range(n):
        th, ph, la = rng.uniform(0, 2 * np.pi, 3)
        t.add(8, [q], gates.u3(th, ph, la))
    t.add(9, [n - 2, 0], gates.CZ)
    t.add(5, [0], np.stack(gates.depolarizing(0.3)))
    return t


def dm_noisy(ref, n, seed, name):
    arrays = dm_tape(n, seed).arrays()
    dm = ref.density_matrix(n)
    outcomes = replay(arrays, dm)
    np.savez(OUT / name, **arrays, n=n, outcomes=np.asarray(outcomes), rho=dm.get_matrix(),
             trace=dm.trace(), purity=dm.purity(), entropy=dm.von_neumann_entropy())


def main():
    orc.build(ref=True)
    ref = orc.CoreLib(orc.REF_LIB)
    sv_u3cx_n8(ref)
    sv_sample_n10(ref)
    sv_fused_n7(ref)
    sv_measure_n5(ref)
    dm_noisy(ref, 4, 21, "dm_noisy_n4.npz")
    dm_noisy(ref, 7, 22, "dm_noisy_n7.npz")
    for p in sorted(OUT.glob("*.npz")):
        print(f"{p.name}: {p.stat().st_size} bytes")


if __name__ == "__main__":
    main()
