"""Task planning, seed tree and aggregation of invariantq_amd/tasks.py on the
CPU, against the semantics of invariantq/distributed/planner.py:110-163 and
aggregation.py:60-120 (hand-derived cases)."""

import hashlib
import struct

import pytest

from invariantq_amd import tasks


def test_seed_tree_known_answers():
    # SHA-256 of the literal bytes, first 8 bytes read as a big-endian unsigned integer
    for master, job, i in [(42, "cfg0", 0), (7, "retry-job", 1), (0, "x", 12)]:
        want = struct.unpack(">Q", hashlib.sha256(b"%d:%s:%d" % (master, job.encode(), i)).digest()[:8])[0]
        assert tasks.task_seed(master, job, i) == want
    assert tasks.task_seed(42, "cfg0", 0) != tasks.task_seed(42, "cfg0", 1)


@pytest.mark.parametrize("total,workers,shots", [
    (10, 4, [3, 3, 2, 2]), (2, 4, [1, 1]), (8, 1, [8]), (20, 2, [10, 10]), (7, 7, [1] * 7)])
def test_plan_split(total, workers, shots):
    plan = tasks.plan_tasks("job", total, 42, workers)
    assert [t.shots for t in plan] == shots
    assert [t.index for t in plan] == list(range(len(shots)))
    assert [t.task_id for t in plan] == [f"job_task_{i}" for i in range(len(shots))]
    assert all(t.seed == tasks.task_seed(42, "job", t.index) for t in plan)
    assert sum(t.shots for t in plan) == total


@pytest.mark.parametrize("args", [("", 10, 42, 4), ("j", 0, 42, 4), ("j", True, 42, 4), ("j", 10, -1, 4),
                                  ("j", 10, 42, 0), (5, 10, 42, 4)])
def test_plan_rejects(args):
    with pytest.raises(ValueError):
        tasks.plan_tasks(*args)


def test_aggregate_parity_and_duplicates():
    r0 = {"task_id": "a", "shots": 4, "counts": {"00": 3, "11": 1}}
    r1 = {"task_id": "b", "shots": 4, "counts": {"01": 2, "10": 2}}
    agg = tasks.aggregate([r0, r1, dict(r1)])
    assert agg["counts"] == {"00": 3, "11": 1, "01": 2, "10": 2}
    assert agg["total_shots"] == 8
    assert agg["duplicate_results_discarded"] == 1
    assert agg["expectation"] == pytest.approx(0.0)   # (3 + 1 - 2 - 2) / 8
    assert agg["standard_error"] == pytest.approx((1.0 / 8) ** 0.5)
    one = tasks.aggregate([r0])
    assert one["expectation"] == pytest.approx(1.0) and one["standard_error"] == 0.0
    with pytest.raises(ValueError):
        tasks.aggregate([{"task_id": "c", "shots": 5, "counts": {"0": 4}}])


def test_counts_of_msb_first():
    assert tasks.counts_of([1, 1, 4], 3) == {"001": 2, "100": 1}
