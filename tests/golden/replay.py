"""Replays a golden op tape through any binding with the reference method names
(apply_1q / apply_2q / apply_fused / measure / reset_qubit for statevectors,
apply_1q_unitary / apply_2q_unitary / apply_1q_kraus / measure / reset_qubit
for density matrices).  Tapes are written by tests/golden/make_golden.py.

Kind codes:
  0 apply_1q   1 apply_2q   2 apply_fused (MSB-first qubit list)
  3 measure(rand) -> outcome   4 reset_qubit(rand)
  5 dm apply_1q_kraus (matrices stacked)   6 dm reset_qubit   7 dm measure(rand) -> outcome
  8 dm apply_1q_unitary   9 dm apply_2q_unitary
"""

from __future__ import annotations

from typing import List

import numpy as np


def tape_ops(tape):
    kinds = tape["op_kind"]
    qubits = tape["op_qubits"]
    off = tape["op_mat_off"]
    mats = tape["mats"]
    rand = tape["op_rand"]
    for i, kind in enumerate(kinds):
        qs = [int(q) for q in qubits[i] if q >= 0]
        m = mats[off[i]:off[i + 1]]
        yield int(kind), qs, m, float(rand[i])


def replay(tape, target, reverse_fused_for_lsb_first: bool = False) -> List[int]:
    """Apply the tape to `target`; returns measurement outcomes in order.

    With reverse_fused_for_lsb_first, k=3,4 fused gates get their qubit list
    reversed: the reference native core indexes k=3,4 blocks LSB-first
    (core/src/statevector.cpp:133-139), which equals MSB-first on the reversed list.
    """
    outcomes = []
    for kind, qs, m, r in tape_ops(tape):
        if kind == 0:
            target.apply_1q(qs[0], m.reshape(2, 2))
        elif kind == 1:
            target.apply_2q(qs[0], qs[1], m.reshape(4, 4))
        elif kind == 2:
            d = 1 << len(qs)
            q = list(reversed(qs)) if (reverse_fused_for_lsb_first and len(qs) >= 3) else qs
            target.apply_fused(q, m.reshape(d, d))
        elif kind == 3:
            outcomes.append(int(target.measure(qs[0], r)))
        elif kind == 4:
            target.reset_qubit(qs[0], r)
        elif kind == 5:
            target.apply_1q_kraus(qs[0], list(m.reshape(-1, 2, 2)))
        elif kind == 6:
            target.reset_qubit(qs[0])
        elif kind == 7:
            outcomes.append(int(target.measure(qs[0], r)))
        elif kind == 8:
            target.apply_1q_unitary(qs[0], m.reshape(2, 2))
        elif kind == 9:
            target.apply_2q_unitary(qs[0], qs[1], m.reshape(4, 4))
        else:
            raise ValueError(f"unknown tape op kind {kind}")
    return outcomes
