"""Seed-tree sampling tasks from one GPU state: every task's indices equal a
separate sample(shots, seed) call on the GPU and the oracle's
restatement of statevector.cpp:235-256 with that seed."""

import numpy as np
import pytest

from invariantq_amd import tasks, workloads

pytestmark = pytest.mark.gpu


def test_tasks_match_single_calls_and_oracle(oracle_lib):
    from invariantq_amd.simulator import StateVectorSimulator
    n = 16
    ops = workloads.random_u3_cx(n, 8, seed=31)
    sim = StateVectorSimulator()
    sv, _ = sim.evolve(n, ops)
    plan = tasks.plan_tasks("gpu-job", 10007, 42, 5)
    print(f"{len(plan)} tasks, n={n}", flush=True)
    got = sv.sample_tasks([t.seed for t in plan], [t.shots for t in plan])
    ref = oracle_lib.statevector(n)
    for op in ops:
        (ref.apply_1q(op.qubits[0], op.matrix) if len(op.qubits) == 1 else ref.apply_2q(op.qubits[0], op.qubits[1], op.matrix))
    cdf_ref = np.cumsum(ref.get_probabilities())
    for t, g in zip(plan, got):
        assert g.size == t.shots
        assert np.array_equal(g, sv.sample(t.shots, t.seed))
        u = oracle_lib.uniforms(t.seed, t.shots)
        j = np.searchsorted(cdf_ref, u)
        near = np.zeros(u.size, dtype=bool)
        for off in (-1, 0):
            jj = np.clip(j + off, 0, cdf_ref.size - 1)
            near |= np.abs(cdf_ref[jj] - u) < 1e-12
        assert not near.any()
        assert np.array_equal(g, ref.sample(t.shots, t.seed))


def test_run_job_aggregate():
    from invariantq_amd.simulator import StateVectorSimulator
    ghz = [("gate", "h", [0], None), ("gate", "cx", [0, 1], None), ("gate", "cx", [1, 2], None)]
    out = tasks.run_job(StateVectorSimulator(), 3, ghz, "ghz", 1001, master_seed=7, num_workers=3)
    assert out["total_shots"] == 1001
    assert set(out["counts"]) == {"000", "111"}
    assert [t["shots"] for t in out["tasks"]] == [334, 334, 333]
    # GHZ-3 parity: "000" even, "111" odd
    p0 = out["counts"]["000"] / 1001
    assert out["expectation"] == pytest.approx(2 * p0 - 1)


def test_sample_tasks_rejects_bad_lengths():
    from invariantq_amd.core.binding import NativeStateVector
    sv = NativeStateVector(4)
    with pytest.raises(ValueError):
        sv.sample_tasks([1, 2], [3])
    with pytest.raises(ValueError):
        sv.sample_tasks([], [])
