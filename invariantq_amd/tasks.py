"""Seed-tree sampling tasks served from one simulation.

The reference splits a job's shots into tasks, gives each task a seed from a
SHA-256 seed tree (invariantq/distributed/planner.py:110-163), and has every
worker re-simulate the whole circuit to draw its shots
(invariantq/distributed/worker.py:52-95).  The sampled indices depend only
on the final state and each task's seed, so here the circuit is simulated
once, its CDF built once on the GPU, and every task's shots are drawn in one
call (``iq_statevector_sample_tasks``): task t gets exactly the indices a
worker running ``sample(shots_t, seed_t)`` on that state would.  Results
are merged the way invariantq/distributed/aggregation.py:60-120 merges task
results.
"""

from __future__ import annotations

import hashlib
import math
from dataclasses import dataclass
from typing import Dict, List, Sequence

import numpy as np


def task_seed(master_seed: int, job_id: str, partition_index: int) -> int:
    """First 8 bytes, big-endian, of SHA-256(f"{master}:{job}:{index}")
    (invariantq/distributed/planner.py:110-115)."""
    digest = hashlib.sha256(f"{master_seed}:{job_id}:{partition_index}".encode("utf-8")).digest()
    return int.from_bytes(digest[:8], byteorder="big")


@dataclass(frozen=True)
class Task:
    task_id: str
    job_id: str
    index: int
    shots: int
    seed: int


def plan_tasks(job_id: str, total_shots: int, master_seed: int = 42, num_workers: int = 4) -> List[Task]:
    """The task split of planner.plan_job (planner.py:117-163): shots divided
    evenly, the first `remainder` tasks one more, tasks with no shots left out."""
    if not isinstance(job_id, str) or not job_id:
        raise ValueError("job_id must be a non-empty string")
    if isinstance(total_shots, bool) or not isinstance(total_shots, int) or total_shots <= 0:
        raise ValueError("total_shots must be a positive integer")
    if isinstance(master_seed, bool) or not isinstance(master_seed, int) or master_seed < 0:
        raise ValueError("master_seed must be a non-negative integer")
    if isinstance(num_workers, bool) or not isinstance(num_workers, int) or num_workers <= 0:
        raise ValueError("num_workers must be a positive integer")
    per, remainder = divmod(total_shots, num_workers)
    tasks = []
    for i in range(num_workers):
        shots = per + (1 if i < remainder else 0)
        if shots == 0:
            continue
        tasks.append(Task(f"{job_id}_task_{i}", job_id, i, shots, task_seed(master_seed, job_id, i)))
    return tasks


def counts_of(samples, n: int) -> Dict[str, int]:
    """MSB-first bit strings, as the simulator's counts."""
    idx, cnt = np.unique(np.asarray(samples, dtype=np.uint64), return_counts=True)
    return {f"{int(i):0{n}b}": int(c) for i, c in zip(idx, cnt)}


def sample_tasks(state, tasks: Sequence[Task], n: int) -> List[Dict]:
    """Every task's counts from one engine state in one call."""
    per_task = state.sample_tasks([t.seed for t in tasks], [t.shots for t in tasks])
    return [{"task_id": t.task_id, "job_id": t.job_id, "shots": t.shots, "seed": t.seed,
             "counts": counts_of(s, n), "samples": s} for t, s in zip(tasks, per_task)]


def aggregate(results: Sequence[Dict]) -> Dict:
    """Merged counts, probabilities, parity expectation <Z...Z> with its
    standard error and 95% interval (aggregation.py:60-120); a repeated
    task_id is counted once and reported as a duplicate."""
    total: Dict[str, int] = {}
    seen, shots, dup = set(), 0, 0
    for r in results:
        if r["task_id"] in seen:
            dup += 1
            continue
        if sum(r["counts"].values()) != r["shots"]:
            raise ValueError(f"result {r['task_id']!r} counts do not sum to its declared shots")
        seen.add(r["task_id"])
        shots += r["shots"]
        for b, c in r["counts"].items():
            total[b] = total.get(b, 0) + c
    probs = {b: c / shots for b, c in total.items()} if shots else {}
    exp = sum((-1.0 if b.count("1") % 2 else 1.0) * p for b, p in probs.items())
    var = max(0.0, 1.0 - exp ** 2)
    se = math.sqrt(var / shots) if shots else 0.0
    return {"counts": total, "probabilities": probs, "total_shots": shots, "expectation": exp,
            "variance": var, "standard_error": se, "confidence_interval": (exp - 1.96 * se, exp + 1.96 * se),
            "completed_tasks": sorted(seen), "duplicate_results_discarded": dup}


def run_job(simulator, num_qubits: int, ops: Sequence, job_id: str, total_shots: int, master_seed: int = 42,
            num_workers: int = 4) -> Dict:
    """Plan, simulate once (``simulator.evolve``), sample every task in one
    call, aggregate.  Returns the aggregate plus the per-task results."""
    tasks = plan_tasks(job_id, total_shots, master_seed, num_workers)
    state, _ = simulator.evolve(num_qubits, ops)
    results = sample_tasks(state, tasks, num_qubits)
    out = aggregate(results)
    out["tasks"] = results
    return out
