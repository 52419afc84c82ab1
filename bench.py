#!/usr/bin/env python3
"""Benchmark of the gfx950 statevector hot path, every workload in one line.

Default (N=1): the headline is the fused-gate workload -- a 30-qubit
complex128 state, 200 seeded Haar-random 32x32 unitaries on random 5-qubit
subsets (even blocks on qubits 0-9, odd blocks unrestricted), applied through
the C ABI (iq_statevector_apply_fused).  One step = one pass of all 200
blocks.  value = achieved HBM TB/s over the timed region with algorithmic
bytes 2 * 2^30 * 16 B per block.  The same JSON line carries, each with its
own roofline (dominant kernel, algorithmic bytes per launch, HIP-event
durations on the stream the kernel runs on), error statement and CPU
baseline (the oracle timed on a bounded sample on the host cores):
  complex64   the same fused pass with complex64 amplitudes
  circuit20   20 qubits, 20 layers of U3 + brick CX, 10^4 seed-tree shots
  adjoint24   24-qubit, 576-parameter ansatz: adjoint vs GPU parameter shift
  batch       4096 ansatz circuits (n 10..20, 4..16 layers) in one launch
  dm14        14-qubit noisy density matrix (Kraus-apply TB/s)

    python bench.py [--steps K] [--warmup W] [--no-cpu] [--gpus 1]
    python bench.py --workload circuit20|dm14|batch4096|adjoint24|bandwidth|primitives
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_TBPS = 8.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_threads() -> int:
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def host_info() -> dict:
    """The host the CPU baselines run on: model, NUMA nodes, threads used."""
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        numa = len([d for d in Path("/sys/devices/system/node").iterdir() if d.name.startswith("node")])
    except OSError:
        numa = None
    return {"cpu_model": model, "numa_nodes": numa, "host_cpus": os.cpu_count()}


def cpu_leg(value, unit, kind, sample) -> dict:
    out = {"value": value, "unit": unit, "cores": cpu_threads(), "kind": kind, "sample": sample}
    out.update(host_info())
    return out


def roofline(kernel: str, bytes_per_launch: float, avg_s: float, launches: int, traffic=None, bound="hbm",
             peak_gbps: float = None) -> dict:
    """Roofline of one kernel: algorithmic bytes per launch over its average
    HIP-event duration, against MI355X's 8 TB/s HBM peak."""
    peak = peak_gbps if peak_gbps is not None else HBM_PEAK_TBPS * 1000
    ach = bytes_per_launch / avg_s / 1e9
    return {"bound": bound, "achieved": round(ach, 1), "peak": peak, "unit": "GB/s", "frac": round(ach / peak, 4),
            "traffic": traffic, "kernel": kernel, "algorithmic_bytes_per_launch": int(bytes_per_launch),
            "avg_ms": round(avg_s * 1e3, 4), "launches_timed": launches}


def fused_workload(n: int, blocks: int, seed: int = 20260):
    from invariantq_amd import workloads
    return workloads.u3_layer(n, seed=seed), workloads.fused_blocks(n, blocks, seed=seed + 1)


def time_fused(binding, torch, n: int, precision: str, steps: int, warmup: int, blocks: int):
    """Runs the fused-gate pass; returns wall time per step and per-kernel HIP-event durations."""
    init, ops = fused_workload(n, blocks)
    sv = binding.NativeStateVector(n, precision=precision)
    for op in init:
        sv.apply_1q(op.qubits[0], op.matrix)
    stream = torch.cuda.ExternalStream(sv.stream)
    mats = [np.asarray(op.matrix) for op in ops]
    qubits = [list(op.qubits) for op in ops]
    targets = [sorted(q) for q in qubits]
    for _ in range(warmup):
        for q, m in zip(qubits, mats):
            sv.apply_fused(q, m)
    sv.synchronize()
    events = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in ops]
    per_kernel = [[] for _ in ops]
    t0 = time.perf_counter()
    for _ in range(steps):
        for i, (q, m) in enumerate(zip(qubits, mats)):
            events[i][0].record(stream)
            sv.apply_fused(q, m)
            events[i][1].record(stream)
        sv.synchronize()
        for i, (a, b) in enumerate(events):
            per_kernel[i].append(a.elapsed_time(b) * 1e-3)
    wall = (time.perf_counter() - t0) / steps
    norm = sv.norm_squared()
    # Calibration for HBM-counter runs: one in-place scale pass of known size
    # (reads and writes 2^n amplitudes once), launched after the timed region.
    sv.normalize()
    sv.synchronize()
    del sv
    return wall, per_kernel, targets, norm


def measured_traffic(name: str):
    """HBM bytes per launch of a dominant kernel from a committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (profiles/<name>_traffic.json), or None."""
    p = ROOT / "profiles" / f"{name}_traffic.json"
    if not p.exists():
        return None
    try:
        return json.loads(p.read_text()).get("hbm_bytes_per_launch")
    except (OSError, ValueError):
        return None


def _chunked_diff(re_a, im_a, re_b, im_b, chunk=1 << 26):
    """max |a_i - b_i| and |<a|b>|^2 / (|a|^2 |b|^2) over split re/im arrays."""
    mx, dot, na, nb = 0.0, 0.0 + 0.0j, 0.0, 0.0
    for i in range(0, re_a.size, chunk):
        ar, ai, br, bi = re_a[i:i + chunk], im_a[i:i + chunk], re_b[i:i + chunk], im_b[i:i + chunk]
        mx = max(mx, float(np.max(np.hypot(ar - br, ai - bi))))
        dot += complex(np.dot(ar, br) + np.dot(ai, bi), np.dot(ar, bi) - np.dot(ai, br))
        na += float(np.dot(ar, ar) + np.dot(ai, ai))
        nb += float(np.dot(br, br) + np.dot(bi, bi))
    return mx, abs(dot) ** 2 / (na * nb)


def cpu_fused_baseline(binding, n: int, blocks_to_time: int):
    """The oracle's fused loop (reference statevector.cpp:89-152, k<=4 guard
    lifted) timed on the first blocks of the same workload on the host cores,
    and the error statement: the GPU's complex128 state after the same U3
    layer and blocks against the oracle's, element by element."""
    from oracle import oracle as orc
    init, ops = fused_workload(n, 200)
    core = orc.oracle()
    ref = core.statevector(n)
    for op in init:
        ref.apply_1q(op.qubits[0], op.matrix)
    t = 0.0
    for op in ops[:blocks_to_time]:
        t0 = time.perf_counter()
        ref.apply_fused(op.qubits, op.matrix)
        t += time.perf_counter() - t0
    bytes_per_block = 2 * (1 << n) * 16
    out = cpu_leg(round(bytes_per_block * blocks_to_time / t / 1e12, 5), "TB/s", "port",
                  f"oracle C++/OpenMP fused loop on the first {blocks_to_time} of the 200 blocks at {n} qubits "
                  f"complex128 ({t:.1f} s); per-block time {t / blocks_to_time:.2f} s")
    sv = binding.NativeStateVector(n)
    for op in init:
        sv.apply_1q(op.qubits[0], op.matrix)
    for op in ops[:blocks_to_time]:
        sv.apply_fused(op.qubits, op.matrix)
    N = 1 << n
    rg, ig, ro, io = (np.empty(N) for _ in range(4))
    binding.lib().iq_statevector_get_amplitudes(sv._ptr, binding._dp(rg), binding._dp(ig))
    binding.check_error()
    del sv
    core.lib.iq_statevector_get_amplitudes(ref._ptr, binding._dp(ro), binding._dp(io))
    del ref
    mx, fid = _chunked_diff(rg, ig, ro, io)
    err = {"what": f"GPU complex128 state vs the oracle after the U3 layer and the first {blocks_to_time} blocks "
                   f"at {n} qubits", "max_abs_amp_diff": mx, "one_minus_fidelity": 1.0 - fid,
           "bar": "max |d amp| <= 1e-12"}
    return out, err


class _DeviceArray:
    """A raw device buffer described by __cuda_array_interface__, for torch.as_tensor."""

    def __init__(self, ptr: int, count: int, typestr: str):
        self.__cuda_array_interface__ = {"shape": (count,), "typestr": typestr, "data": (ptr, False),
                                         "version": 2, "strides": None}


def c64_fidelity(binding, torch, n: int, blocks: int):
    """One pass of the U3 layer and the blocks in complex64 and in complex128;
    the overlap of the two states on the device (FP64 sums)."""
    init, ops = fused_workload(n, blocks)
    states = {}
    for prec in ("complex128", "complex64"):
        sv = binding.NativeStateVector(n, precision=prec)
        for op in init:
            sv.apply_1q(op.qubits[0], op.matrix)
        for op in ops:
            sv.apply_fused(op.qubits, op.matrix)
        sv.synchronize()
        states[prec] = sv
    N = 1 << n
    a = torch.as_tensor(_DeviceArray(states["complex64"].device_ptr, N, "<c8"), device="cuda")
    b = torch.as_tensor(_DeviceArray(states["complex128"].device_ptr, N, "<c16"), device="cuda")
    dot, na, nb, mx = 0.0 + 0.0j, 0.0, 0.0, 0.0
    chunk = 1 << 26
    for i in range(0, N, chunk):
        x = a[i:i + chunk].to(torch.complex128)
        y = b[i:i + chunk]
        dot += complex(torch.vdot(x, y).item())
        na += float(torch.vdot(x, x).real.item())
        nb += float(torch.vdot(y, y).real.item())
        mx = max(mx, float((x - y).abs().max().item()))
    torch.cuda.synchronize()
    del a, b, states
    return {"what": f"complex64 vs complex128 after one pass (U3 layer + {blocks} blocks) at {n} qubits",
            "one_minus_fidelity": 1.0 - abs(dot) ** 2 / (na * nb), "max_abs_amp_diff": mx,
            "bar": "fidelity >= 1 - 1e-6, max |d amp| <= 2e-5"}


def adjoint_vs_shift(torch, n: int, layers: int, steps: int, shift_params: int, cpu: bool = False):
    """Adjoint gradient (complex64 state, FP64 accumulation) vs parameter-shift on the same GPU."""
    from invariantq_amd import adjoint, workloads
    ans = workloads.Ansatz(n, layers)
    terms = workloads.tfim_terms(n)
    rng = np.random.default_rng(2026)
    params = rng.uniform(0.0, 2 * np.pi, ans.num_params)
    fast = adjoint.GpuCircuit.from_ansatz(ans, "complex64")
    fast.set_observable(terms)
    for _ in range(2):  # warm-up (compiles the tape, allocates the partial buffers and stored states)
        fast.gradient(params)
    t0 = time.perf_counter()
    for _ in range(steps):
        e64, g64 = fast.gradient(params)
    t_adj = (time.perf_counter() - t0) / steps
    which = list(np.linspace(0, ans.num_params - 1, shift_params).astype(int))
    fast.parameter_shift(params, which[:2])  # warm-up
    t0 = time.perf_counter()
    ps64 = fast.parameter_shift(params, which)
    t_ps = (time.perf_counter() - t0) / len(which)
    exact = adjoint.GpuCircuit.from_ansatz(ans, "complex128")
    exact.set_observable(terms)
    e128, g128 = exact.gradient(params)
    ps128 = exact.parameter_shift(params, which)
    rel_adj = float(np.linalg.norm(g64 - g128) / np.linalg.norm(g128))
    rel_adj_vs_ps = float(np.linalg.norm(g64[which] - ps128) / np.linalg.norm(ps128))
    rel_ps = float(np.linalg.norm(ps64 - ps128) / np.linalg.norm(ps128))
    # one more gradient with every launch bracketed by events on the circuit's
    # stream: the reverse tile passes are the dominant kernel
    fast.set_timing(True)
    fast.gradient(params)
    times = fast.launch_times()
    fast.set_timing(False)
    n_amp = 1 << n
    segs = fast.num_segments
    # HBM traffic of the tiled sweeps: a forward pass reads and writes the
    # state (2 N s), the observable's three tile passes move 8 N s, a reverse
    # pass reads the stored psi and reads and writes lambda (3 N s)
    adj_bytes = segs * 2 * n_amp * 8 + 8 * n_amp * 8 + segs * 3 * n_amp * 8
    # the same gradient gate by gate: each gate a forward pass, each gate a
    # reverse pass over psi and lambda, one observable pass
    gate_bytes = fast.num_gates * 2 * n_amp * 8 + 2 * n_amp * 8 + fast.num_gates * 4 * n_amp * 8
    by_kind = {}
    for k, ms in times:
        by_kind.setdefault(k, []).append(ms * 1e-3)
    rev, fwd = by_kind.get("reverse", []), by_kind.get("forward", [])
    launch_stats = {k: {"launches": len(v), "total_ms": round(sum(v) * 1e3, 4), "avg_ms": round(np.mean(v) * 1e3, 4)}
                    for k, v in by_kind.items()}
    if fwd:
        launch_stats["forward"]["achieved_TBps"] = round(2 * n_amp * 8 / np.mean(fwd) / 1e12, 3)
    out_roof = None
    if rev:
        # a reverse pass reads the stored psi and reads and writes lambda: 3 N 8 B
        out_roof = roofline("k_tile_rev_lam<float> (reverse tile pass)", 3 * n_amp * 8, float(np.mean(rev)), len(rev),
                            traffic=measured_traffic("adjoint_24q_c64") if n == 24 else None)
        out_roof["per_pass_in_layer_TBps"] = [round(3 * n_amp * 8 / float(np.mean(rev[i::3])) / 1e12, 3)
                                              for i in range(min(3, len(rev)))]
    cpu_out = None
    if cpu:
        # the oracle's FP64 parameter shift: two full evaluations per parameter;
        # timed on one layer of the circuit and the observable, extrapolated
        from oracle import oracle as orc
        core = orc.oracle()
        ops = orc.hea_ops(n, layers, params)
        per_layer = len(ops) // layers
        ref = core.statevector(n)
        t0 = time.perf_counter()
        for op in ops[:per_layer]:
            qs, m = op
            (ref.apply_1q(qs[0], m) if len(qs) == 1 else ref.apply_2q(qs[0], qs[1], m))
        t_layer = time.perf_counter() - t0
        t0 = time.perf_counter()
        for c_, o_, q_ in orc.tfim_terms(n):
            ref.expectation_value_pauli(o_, q_)
        t_obs = time.perf_counter() - t0
        del ref
        t_eval = layers * t_layer + t_obs
        cpu_out = cpu_leg(round(1.0 / (2.0 * t_eval), 5), "gradients/s", "port",
                          f"oracle FP64 parameter shift (2 evaluations per parameter); one evaluation "
                          f"= {layers} x one timed layer ({t_layer:.2f} s) + the {len(orc.tfim_terms(n))}-term "
                          f"observable ({t_obs:.2f} s) at {n} qubits")
    return {
        "workload": f"{n} qubits, {layers} layers RY+RZ+CZ ring ({fast.num_gates} gates, {ans.num_params} params), "
                    f"TFIM open chain, complex64 state, FP64 accumulation",
        "adjoint_s_per_gradient": round(t_adj, 5),
        "adjoint_grads_per_s": round(ans.num_params / t_adj, 1),
        "param_shift_s_per_param": round(t_ps, 6),
        "param_shift_grads_per_s": round(1.0 / t_ps, 2),
        "param_shift_timed_params": len(which),
        "speedup_adjoint_vs_param_shift": round(t_ps * ans.num_params / t_adj, 1),
        "rel_err_adjoint_c64_vs_adjoint_c128": rel_adj,
        "rel_err_adjoint_c64_vs_param_shift_c128": rel_adj_vs_ps,
        "rel_err_param_shift_c64_vs_c128": rel_ps,
        "energy_c64": e64, "energy_c128": e128,
        "tile_passes_per_sweep": segs,
        "adjoint_hbm_GB_per_gradient": round(adj_bytes / 1e9, 2),
        "adjoint_effective_TBps": round(adj_bytes / t_adj / 1e12, 3),
        "gate_by_gate_GB_per_gradient": round(gate_bytes / 1e9, 1),
        "gate_by_gate_equivalent_TBps": round(gate_bytes / t_adj / 1e12, 2),
        "launches": launch_stats,
        "roofline": out_roof,
        "error": {"what": "complex64 adjoint gradient vs complex128 parameter shift (timed parameters) and "
                          "vs the complex128 adjoint (all)",
                  "rel_l2_vs_param_shift_c128": rel_adj_vs_ps, "rel_l2_vs_adjoint_c128": rel_adj,
                  "bar": "relative L2 <= 1e-4"},
        "cpu_baseline": cpu_out,
    }


def circuit_workload(binding, torch, steps: int, warmup: int, cpu: bool):
    """20-qubit complex128 random circuit (20 layers of U3 on every qubit + brick CX,
    590 gates) through the C ABI, then 10^4 shots from the seed-tree seed."""
    import hashlib
    from invariantq_amd import workloads
    n, shots = 20, 10000
    ops = workloads.random_u3_cx(n, 20, seed=2024)
    seed = int.from_bytes(hashlib.sha256(b"42:cfg0:0").digest()[:8], "big")
    sv = binding.NativeStateVector(n)
    stream = torch.cuda.ExternalStream(sv.stream)

    def run():
        sv.reset()
        for op in ops:
            if len(op.qubits) == 1:
                sv.apply_1q(op.qubits[0], op.matrix)
            else:
                sv.apply_2q(op.qubits[0], op.qubits[1], op.matrix)

    for _ in range(warmup):
        run()
        sv.sample(shots, seed)
    sv.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    t_gates, t_total, t_gpu = 0.0, 0.0, 0.0
    for _ in range(steps):
        t0 = time.perf_counter()
        ev[0].record(stream)
        run()
        ev[1].record(stream)
        sv.synchronize()
        t1 = time.perf_counter()
        samples = sv.sample(shots, seed)
        t2 = time.perf_counter()
        t_gates += t1 - t0
        t_total += t2 - t0
        t_gpu += ev[0].elapsed_time(ev[1]) * 1e-3
    t_gates, t_total, t_gpu = t_gates / steps, t_total / steps, t_gpu / steps
    # one more run with every gate bracketed by events on the handle's stream:
    # per-kind launch durations (the one-qubit U3 launches dominate)
    pairs = []
    sv.reset()
    for op in ops:
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        if len(op.qubits) == 1:
            sv.apply_1q(op.qubits[0], op.matrix)
        else:
            sv.apply_2q(op.qubits[0], op.qubits[1], op.matrix)
        b.record(stream)
        pairs.append((len(op.qubits), a, b))
    sv.synchronize()
    d1 = [a.elapsed_time(b) * 1e-3 for k, a, b in pairs if k == 1]
    d2 = [a.elapsed_time(b) * 1e-3 for k, a, b in pairs if k == 2]
    gate_bytes = 2 * (1 << n) * 16
    out = {"workload": "20 qubits complex128, 20 layers U3 + brick CX (590 gates), 10^4 seed-tree shots",
           "gate_applications_per_s": round(len(ops) / t_gates, 1),
           "gate_applications_per_s_gpu_time": round(len(ops) / t_gpu, 1),
           "circuit_ms": round(t_gates * 1e3, 3), "circuit_gpu_ms": round(t_gpu * 1e3, 3),
           "circuit_plus_shots_ms": round(t_total * 1e3, 3),
           "effective_TBps_gpu_time": round(len(ops) * 2 * (1 << n) * 16 / t_gpu / 1e12, 3),
           "first_samples": [int(x) for x in samples[:4]],
           # the 16 MiB state stays in the 256 MiB Infinity Cache: the HBM roof is
           # a reference line here, launch overhead and cache bandwidth bound it
           "roofline": roofline("U3 one-qubit gate launches (k_gate_direct_val / k_gate_lanes)", gate_bytes,
                                float(np.mean(d1)), len(d1)),
           "cx_launch_avg_ms": round(float(np.mean(d2)) * 1e3, 4),
           "cx_launch_TBps": round(gate_bytes / float(np.mean(d2)) / 1e12, 3)}
    # the same circuit fused into blocks of up to 5 qubits (fusion.py)
    from invariantq_amd import fusion
    t0 = time.perf_counter()
    fused = fusion.fuse_statevector(ops, 5)
    t_fuse = time.perf_counter() - t0
    unfused_amps = sv.get_amplitudes()

    seq = [(f.qubits, f.matrix) for f in fused]

    def run_fused():
        sv.reset()
        sv.apply_fused_sequence(seq)

    for _ in range(warmup):
        run_fused()
    sv.synchronize()
    tf, tf_gpu = 0.0, 0.0
    for _ in range(steps):
        t0 = time.perf_counter()
        ev[0].record(stream)
        run_fused()
        ev[1].record(stream)
        sv.synchronize()
        tf += time.perf_counter() - t0
        tf_gpu += ev[0].elapsed_time(ev[1]) * 1e-3
    tf, tf_gpu = tf / steps, tf_gpu / steps
    fs = sv.sample(shots, seed)
    out["fused"] = {"passes": len(fused), "widths": {str(k): sum(len(f.qubits) == k for f in fused) for k in range(1, 6)},
                    "fusion_host_ms": round(t_fuse * 1e3, 2), "circuit_ms": round(tf * 1e3, 3),
                    "circuit_gpu_ms": round(tf_gpu * 1e3, 3), "speedup_gpu_time": round(t_gpu / tf_gpu, 2),
                    "max_abs_amp_diff_vs_unfused": float(np.max(np.abs(sv.get_amplitudes() - unfused_amps))),
                    "samples_equal_unfused": bool(np.array_equal(fs, samples))}
    # the same circuit through the circuit engine's tile passes (CX as H CZ H:
    # each layer's one-qubit runs and CZ signs in about two passes), the state
    # copied into the statevector handle; timed from the first launch to the
    # copy, wall and GPU (events on the circuit's stream, then the copy)
    from invariantq_amd import adjoint
    circ = adjoint.GpuCircuit.from_gates(n, [(op.qubits, op.matrix) for op in ops], "complex128")
    cstream = torch.cuda.ExternalStream(circ.stream)

    def run_tiled():
        circ.run()
        circ.copy_state_to(sv)

    for _ in range(warmup + 1):
        run_tiled()
    sv.synchronize()
    tt, tt_gpu = 0.0, 0.0
    for _ in range(steps):
        t0 = time.perf_counter()
        ev[0].record(cstream)
        circ.run()
        ev[1].record(cstream)
        circ.copy_state_to(sv)
        sv.synchronize()
        tt += time.perf_counter() - t0
        tt_gpu += ev[0].elapsed_time(ev[1]) * 1e-3
    tt, tt_gpu = tt / steps, tt_gpu / steps
    ts_ = sv.sample(shots, seed)
    out["tiled"] = {"passes": int(circ.num_segments), "circuit_ms": round(tt * 1e3, 3),
                    "circuit_gpu_ms": round(tt_gpu * 1e3, 3),
                    "gate_applications_per_s": round(len(ops) / tt, 1),
                    "speedup_vs_gate_by_gate_wall": round(t_gates / tt, 2),
                    "max_abs_amp_diff_vs_unfused": float(np.max(np.abs(sv.get_amplitudes() - unfused_amps))),
                    "samples_equal_unfused": bool(np.array_equal(ts_, samples))}
    del circ
    # seed-tree tasks (4 workers, job "cfg0", master 42): one call on one state,
    # against the reference's way, a re-simulation plus sample() per task
    from invariantq_amd import tasks as T
    plan = T.plan_tasks("cfg0", shots, 42, 4)
    run()
    sv.synchronize()
    t0 = time.perf_counter()
    per = sv.sample_tasks([t.seed for t in plan], [t.shots for t in plan])
    t_one = time.perf_counter() - t0
    t0 = time.perf_counter()
    resim = []
    for t in plan:
        run()
        resim.append(sv.sample(t.shots, t.seed))
    t_re = time.perf_counter() - t0
    out["tasks"] = {"tasks": len(plan), "shots": shots, "one_call_ms": round(t_one * 1e3, 3),
                    "resimulate_per_task_ms": round(t_re * 1e3, 3), "speedup": round(t_re / t_one, 1),
                    "indices_equal": bool(all(np.array_equal(a, b) for a, b in zip(per, resim)))}
    # entanglement entropy: the half cut StateVectorSimulator.run computes (k = 10),
    # and k = 8, which the CPU oracle finishes in seconds
    # (median of 3 calls after the first; the first call also grows the work
    # buffer and loads the kernels, reported on its own)
    sv.synchronize()

    def timed_entropy(qs):
        ts, h = [], None
        for _ in range(4):
            t0 = time.perf_counter()
            h = sv.von_neumann_entropy(qs)
            ts.append(time.perf_counter() - t0)
        return h, ts[0], float(np.median(ts[1:]))

    h_half, t_half_first, t_half = timed_entropy(list(range(n // 2)))
    h8, _, t8 = timed_entropy(list(range(8)))
    out["entropy"] = {"half_cut_qubits": n // 2, "half_cut_bits": h_half, "half_cut_ms": round(t_half * 1e3, 3),
                      "half_cut_first_call_ms": round(t_half_first * 1e3, 3),
                      "k8_bits": h8, "k8_ms": round(t8 * 1e3, 3)}
    if cpu:
        from oracle import oracle as orc
        core = orc.oracle()
        ref = core.statevector(n)
        t0 = time.perf_counter()
        for op in ops:
            (ref.apply_1q(op.qubits[0], op.matrix) if len(op.qubits) == 1
             else ref.apply_2q(op.qubits[0], op.qubits[1], op.matrix))
        t1 = time.perf_counter()
        rs = ref.sample(shots, seed)
        t2 = time.perf_counter()
        a, b = sv.get_amplitudes(), ref.get_amplitudes()
        out["max_abs_amp_err_vs_oracle"] = float(np.max(np.abs(a - b)))
        te = time.perf_counter()
        rh8 = ref.von_neumann_entropy(list(range(8)))
        out["entropy"]["k8_cpu_oracle_ms"] = round((time.perf_counter() - te) * 1e3, 1)
        out["entropy"]["k8_abs_diff_vs_oracle"] = abs(rh8 - h8)
        out["samples_bit_exact_vs_oracle"] = bool(np.array_equal(samples, rs))
        out["error"] = {"what": "final 20-qubit state vs the oracle; 10^4 shots vs the oracle's",
                        "max_abs_amp_diff": out["max_abs_amp_err_vs_oracle"],
                        "samples_bit_exact": out["samples_bit_exact_vs_oracle"], "bar": "max |d amp| <= 1e-12"}
        out["cpu_baseline"] = cpu_leg(round(len(ops) / (t1 - t0), 1), "gate-applications/s", "port",
                                      f"oracle C++/OpenMP, the full 590-gate circuit ({(t1 - t0) * 1e3:.1f} ms) "
                                      f"and 10^4 shots ({(t2 - t1) * 1e3:.1f} ms)")
    return out


def bandwidth_workload(binding, torch, reps: int = 5):
    """Plain passes over 16 GiB (a 30-qubit complex128 state) on the same GPU:
    what read-only and read+write streaming reach here, beside the fused pass.
    torch kernels are timed with events on torch's stream; the engine's
    normalize (one read-only norm pass, then one in-place scale pass) with
    wall time after a synchronize."""
    nbytes = 16 * 2**30
    x = torch.empty(2**30, dtype=torch.complex128, device="cuda")
    x.real.fill_(1.0)
    x.imag.fill_(0.5)
    y = torch.empty_like(x)

    def events(f):
        f()
        torch.cuda.synchronize()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(reps):
            f()
        e.record()
        torch.cuda.synchronize()
        return s.elapsed_time(e) / reps / 1e3

    t_mul = events(lambda: x.mul_(1.0000001))
    t_copy = events(lambda: y.copy_(x))
    t_sum = events(lambda: x.real.sum())
    del x, y
    torch.cuda.empty_cache()
    sv = binding.NativeStateVector(30)
    sv.normalize()
    sv.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        sv.normalize()
    sv.synchronize()
    t_norm = (time.perf_counter() - t0) / reps
    del sv
    return {
        "workload": "plain streaming passes over 16 GiB (2^30 complex128 amplitudes)",
        "torch_inplace_mul_TBps": round(2 * nbytes / t_mul / 1e12, 3),
        "torch_copy_TBps": round(2 * nbytes / t_copy / 1e12, 3),
        "torch_sum_real_part_read_TBps": round(nbytes / t_sum / 1e12, 3),
        "engine_normalize_TBps": round(3 * nbytes / t_norm / 1e12, 3),
        "note": "normalize = read-only norm pass + in-place scale pass (3 x 16 GiB); "
                "rocprofv3 splits it per kernel",
    }


def primitives_workload(binding, torch, n: int = 30, reps: int = 3):
    """Each statevector entry point of the C ABI on one 2^n complex128 state
    (16 GiB at 30 qubits), against its algorithmic bytes (N amplitudes of
    16 B): gates read and write every amplitude (2 N 16); reductions read it
    once (N 16); collapse reads the kept half and writes all (1.5 N 16);
    probabilities read the state and write 8 B each; sampling reads the state
    and writes the 8 B CDF.  Enqueued calls are timed with events on the
    handle's stream around `reps` calls; calls that return data with wall time
    (they synchronize).  Exports are timed through the raw C calls into host
    buffers touched beforehand, so they include the device-to-host copy."""
    from invariantq_amd import gates, workloads
    N, s = 1 << n, 16
    lib = binding.lib()
    dp = binding._dp
    sv = binding.NativeStateVector(n)
    layer = workloads.u3_layer(n, seed=31)
    for op in layer:
        sv.apply_1q(op.qubits[0], op.matrix)
    st = torch.cuda.ExternalStream(sv.stream)
    rng = np.random.default_rng(30)
    u2 = workloads.haar_unitary(rng, 4)
    u3m = workloads.haar_unitary(rng, 8)
    u4m = workloads.haar_unitary(rng, 16)

    def dev(f):
        f()
        sv.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(st)
        for _ in range(reps):
            f()
        b.record(st)
        sv.synchronize()
        return a.elapsed_time(b) / reps / 1e3

    def wall(f, k=reps):
        f()
        sv.synchronize()
        t0 = time.perf_counter()
        for _ in range(k):
            f()
        sv.synchronize()  # calls that end with enqueued work (normalize's scale, reset)
        return (time.perf_counter() - t0) / k

    rows = []

    def row(name, ref, nbytes, t, how, host_bytes=0):
        rows.append({"row": name, "reference": ref, "algorithmic_GB": round(nbytes / 1e9, 3),
                     "ms": round(t * 1e3, 3), "TBps": round(nbytes / t / 1e12, 3),
                     "pct_hbm_peak": round(100 * nbytes / t / 8e12, 1), "timed": how})
        if host_bytes:  # bytes delivered into the caller's (pageable) host buffers
            rows[-1]["host_GBps"] = round(host_bytes / t / 1e9, 1)

    g1 = gates.u3(0.3, 1.1, 2.0)
    for t in (0, 4, 10, 29):
        row(f"a2 apply_1q target {t}", "statevector.cpp:26-47", 2 * N * s, dev(lambda: sv.apply_1q(t, g1)), "events")
    for q0, q1 in ((0, 1), (1, 20), (4, 20), (5, 17), (28, 29)):
        row(f"a3 apply_2q dense ({q0},{q1})", "statevector.cpp:49-87", 2 * N * s,
            dev(lambda: sv.apply_2q(q0, q1, u2)), "events")
    row("a3 apply_2q CX (3,20)", "statevector.cpp:49-87", 2 * N * s, dev(lambda: sv.apply_2q(3, 20, gates.CX)), "events")
    # control at bit 20: only the half with the control set is read and written
    row("a3 apply_2q CX (20,3), controlled: half the state", "statevector.cpp:49-87", N * s,
        dev(lambda: sv.apply_2q(20, 3, gates.CX)), "events")
    row("a3 apply_2q CZ (7,25), diagonal", "statevector.cpp:49-87", 2 * N * s,
        dev(lambda: sv.apply_2q(7, 25, gates.CZ)), "events")
    # Toffoli (controls 21 and 9, target 3): blockdiag(I6, X) on the 3 qubits,
    # MSB-first; only the quarter with both controls set is read and written
    toffoli = np.eye(8, dtype=np.complex128)
    toffoli[6:, 6:] = gates.X
    row("a4 apply_fused Toffoli (21,9,3), controlled: a quarter of the state", "statevector.cpp:89-152",
        0.5 * N * s, dev(lambda: sv.apply_fused([21, 9, 3], toffoli)), "events")
    row("a4 apply_fused k=3 (4,12,27)", "statevector.cpp:89-152", 2 * N * s,
        dev(lambda: sv.apply_fused([4, 12, 27], u3m)), "events")
    row("a4 apply_fused k=4 (0,2,15,22), LDS-tiled", "statevector.cpp:89-152", 2 * N * s,
        dev(lambda: sv.apply_fused([0, 2, 15, 22], u4m)), "events")
    row("a4 apply_fused k=4 (5,9,18,26)", "statevector.cpp:89-152", 2 * N * s,
        dev(lambda: sv.apply_fused([5, 9, 18, 26], u4m)), "events")
    row("a5 measure_prob target 17", "statevector.cpp:154-169", N * s, wall(lambda: sv.measure_prob(17, 0)), "wall")
    row("a7 norm_squared", "statevector.cpp:206-214", N * s, wall(lambda: sv.norm_squared()), "wall")
    row("a7 normalize", "statevector.cpp:216-224", 3 * N * s, wall(lambda: sv.normalize()), "wall")
    row("a10 expval_pauli XZY on (2,9,21)", "statevector.cpp:258-291", N * s,
        wall(lambda: sv.expectation_value_pauli("XZY", [2, 9, 21])), "wall")
    terms = workloads.tfim_terms(n)
    row(f"a10 TFIM <H>, {len(terms)} strings in one call (bytes of one read)", "statevector.cpp:258-291", N * s,
        wall(lambda: sv.expectation_value_pauli_sum(terms)), "wall")
    shots = 10000
    # bytes: one read of the state, then one 4096-amplitude chunk and 16 B per shot
    row(f"a9 sample {shots} shots (one read + a chunk per shot)", "statevector.cpp:235-256",
        N * s + shots * (4096 * s + 16), wall(lambda: sv.sample(shots, 42)), "wall")
    host = np.empty(N, dtype=np.float64)
    host.fill(0.0)
    row("a8 get_probabilities (kernel + device-to-host copy)", "api.cpp:85-90, statevector.cpp:226-233",
        N * s + 8 * N, wall(lambda: lib.iq_statevector_get_probabilities(sv._ptr, dp(host)), 1), "wall", 8 * N)
    host2 = np.empty(N, dtype=np.float64)
    host2.fill(0.0)
    row("a11 get_amplitudes (kernel + device-to-host copy)", "api.cpp:74-83", N * s + 16 * N,
        wall(lambda: lib.iq_statevector_get_amplitudes(sv._ptr, dp(host), dp(host2)), 1), "wall", 16 * N)
    del host, host2
    row("a6 measure target 29 (probability + collapse)", "statevector.cpp:171-190", 2.5 * N * s,
        wall(lambda: sv.measure(29, 0.3)), "wall")
    row("a1 reset", "statevector.cpp:9-17, api.cpp:24-28", N * s, wall(lambda: sv.reset()), "wall")
    binding.check_error()
    del sv
    return {"workload": f"statevector entry points on one {n}-qubit complex128 state ({N * s / 2**30:.0f} GiB)",
            "rows": rows}


def density_matrix_workload(binding, torch, cpu: bool):
    """14-qubit density matrix (2^28 complex128 entries): depth-10 U3 + brick CX circuit,
    depolarizing (p=1e-3) on touched qubits after every gate, amplitude (1e-3) and phase
    (5e-4) damping on every qubit per layer.  Reports Kraus-apply TB/s."""
    from invariantq_amd import workloads
    n = 14
    ops = workloads.noisy_dm_circuit(n, depth=10, seed=77)
    dm = binding.NativeDensityMatrix(n)
    stream = torch.cuda.ExternalStream(dm.stream)
    kinds = {"kraus": [], "u1": [], "u2": []}
    pairs = []
    t0 = time.perf_counter()
    for op in ops:
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        if op.kind == "kraus":
            dm.apply_1q_kraus(op.qubits[0], op.kraus)
            k = "kraus"
        elif len(op.qubits) == 1:
            dm.apply_1q_unitary(op.qubits[0], op.matrix)
            k = "u1"
        else:
            dm.apply_2q_unitary(op.qubits[0], op.qubits[1], op.matrix)
            k = "u2"
        b.record(stream)
        pairs.append((k, a, b))
    dm.synchronize()
    wall = time.perf_counter() - t0
    for k, a, b in pairs:
        kinds[k].append(a.elapsed_time(b) * 1e-3)
    byts = 2 * (1 << (2 * n)) * 16
    trace, purity, z = dm.trace(), dm.purity(), dm.expval_z()
    out = {"workload": "14-qubit density matrix, depth-10 U3 + brick CX, depolarizing + amplitude + phase damping",
           "ops": len(ops), "wall_s": round(wall, 4),
           "trace": trace, "purity": purity, "expval_z": [round(float(v), 12) for v in z]}
    for k, v in kinds.items():
        if v:
            out[k] = {"count": len(v), "avg_ms": round(float(np.mean(v)) * 1e3, 4),
                      "TBps": round(byts / float(np.mean(v)) / 1e12, 4)}
    out["kraus_apply_TBps"] = out["kraus"]["TBps"]
    out["kraus_apply_pct_hbm_peak"] = round(100 * out["kraus"]["TBps"] / HBM_PEAK_TBPS, 2)
    # a one-qubit Kraus channel is one superoperator pass over the 2n-qubit vector
    out["roofline"] = roofline("Kraus apply: superoperator pass on rho (k_gate_direct_val / k_gate_lanes)", byts,
                               float(np.mean(kinds["kraus"])), len(kinds["kraus"]))
    # the same circuit as fused superoperators on up to 2 qubits (fusion.py)
    from invariantq_amd import fusion
    t0 = time.perf_counter()
    fused = fusion.fuse_density_matrix(ops, 2)
    t_fuse = time.perf_counter() - t0
    dm.reset()
    dm.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    a.record(stream)
    for f in fused:
        dm.apply_superop(f.qubits, f.matrix)
    b.record(stream)
    dm.synchronize()
    fwall = time.perf_counter() - t0
    fgpu = a.elapsed_time(b) * 1e-3
    ft, fp, fz = dm.trace(), dm.purity(), dm.expval_z()
    out["fused"] = {"passes": len(fused), "fusion_host_ms": round(t_fuse * 1e3, 2), "wall_s": round(fwall, 4),
                    "gpu_s": round(fgpu, 4), "avg_pass_TBps": round(byts * len(fused) / fgpu / 1e12, 4),
                    "speedup_wall": round(wall / fwall, 2),
                    "trace_diff": abs(ft - trace), "purity_diff": abs(fp - purity),
                    "max_expval_z_diff": float(np.max(np.abs(np.asarray(fz) - np.asarray(z))))}
    out["error"] = {"what": "trace and purity, op by op vs fused superoperators; <Z_i> op by op vs fused",
                    "trace_minus_one": trace - 1.0, "purity": purity,
                    "trace_diff_fused": out["fused"]["trace_diff"], "purity_diff_fused": out["fused"]["purity_diff"],
                    "max_expval_z_diff_fused": out["fused"]["max_expval_z_diff"],
                    "bar": "|trace - 1| <= 1e-12; fused vs op by op <= 1e-12"}
    if cpu:
        from invariantq_amd import gates as G
        from oracle import oracle as orc
        ref = orc.oracle().density_matrix(n)
        t0 = time.perf_counter()
        ref.apply_1q_kraus(3, G.depolarizing(1e-3))
        t_k = time.perf_counter() - t0
        out["cpu_baseline"] = cpu_leg(round(byts / t_k / 1e12, 5), "TB/s", "port",
                                      f"oracle restatement of density_matrix.cpp:147-163 as written (4 copies + "
                                      f"left/right multiplies), one depolarizing Kraus apply at 14 qubits: {t_k:.2f} s")
    return out


def batch_workload(torch, count: int, cpu: bool):
    """4096 ansatz circuits (n ~ U{10..20}, layers ~ U{4..16}, complex64, FP64
    accumulation): energies and full adjoint gradients in one persistent launch,
    tile widths and workgroup size from the autotuner."""
    from invariantq_amd import adjoint, batch, workloads
    specs = workloads.batch_specs(count, seed=3)
    runner = batch.BatchRunner("complex64")
    t0 = time.perf_counter()
    tune = batch.autotune(runner, specs, cache=None, per_n=128)
    t_tune = time.perf_counter() - t0
    gw = tune["group_width"]
    # warm-up: the whole batch, so the timed launch is a steady repeat
    runner.run(specs, tile_bits=tune["tile_bits"], block_threads=tune["block_threads"], group_width=gw, solo=False)
    t0 = time.perf_counter()
    e, g, ms = runner.run(specs, tile_bits=tune["tile_bits"], block_threads=tune["block_threads"], group_width=gw,
                          solo=False)
    wall = time.perf_counter() - t0
    nparams = int(sum(2 * n * l for n, l, _ in specs))
    gate_pass_bytes = sum(3 * n * l * (2 + 4) * (1 << n) * 8 for n, l, _ in specs)
    rng = np.random.default_rng(9)
    check = sorted(rng.choice(count, 12, replace=False).tolist(), key=lambda i: specs[i][0])
    errs = []
    for i in check:
        n, l, p = specs[i]
        c = adjoint.GpuCircuit.from_ansatz(workloads.Ansatz(n, l), "complex128")
        c.set_observable(workloads.tfim_terms(n))
        e_ref, g_ref = c.gradient(p)
        errs.append(float(np.linalg.norm(g[i] - g_ref) / np.linalg.norm(g_ref)))
    out = {"workload": f"{count} ansatz circuits, n ~ U{{10..20}}, layers ~ U{{4..16}}, complex64 state, "
                       "FP64 accumulation, TFIM energy + full adjoint gradient each, one persistent launch",
           "circuits": count, "parameters": nparams,
           "kernel_ms": round(ms, 3), "wall_ms": round(wall * 1e3, 3),
           "circuits_per_s": round(count / (ms * 1e-3), 1),
           "gradients_per_s": round(nparams / (ms * 1e-3), 1),
           "gate_pass_equivalent_TBps": round(gate_pass_bytes / (ms * 1e-3) / 1e12, 3),
           "autotune": {"tile_bits": tune["tile_bits"], "block_threads": tune["block_threads"], "group_width": gw,
                        "group_width_ms": tune["timings"].get("group_width"),
                        "threads_ms": tune["timings"].get("block_threads"),
                        "seconds": round(t_tune, 2)},
           "max_rel_grad_err_vs_complex128": max(errs), "checked_circuits": [specs[i][0] for i in check]}
    # per size class: each class launched alone, with the configuration the
    # autotuner picks for that launch (a class alone has fewer circuits than
    # the whole batch, so its best occupancy differs).  Circuits that fit the
    # tile width live in LDS (bound: LDS, 128 B/clk/CU); wider ones stream psi
    # and lambda through HBM in L + 1 tiled passes each way, each pass serving
    # half of two layers (bound: HBM).
    lds_peak = 256 * 128 * 2.4e9
    gm, elem = 4, 8  # complex64: 4 qubits per register group, 8-byte amplitudes
    classes = {}
    launch_bytes = 0.0
    for n in sorted({s[0] for s in specs}):
        cls = [s for s in specs if s[0] == n]
        ct = batch.autotune(runner, cls, cache=None, per_n=64)
        cfg = dict(tile_bits=ct["tile_bits"], block_threads=ct["block_threads"], group_width=ct["group_width"])
        runner.run(cls, solo=False, **cfg)
        _, _, cms = runner.run(cls, solo=False, **cfg)
        b, N = ct["tile_bits"][n], 1 << n
        L = sum(s[1] for s in cls)
        if n <= b:
            sweeps_f, sweeps_r = -(-n // gm) + 1, -(-n // gm) + 1
            byts = L * (sweeps_f * 2 + sweeps_r * 4) * N * elem
            bound, peak = "lds", lds_peak
        else:
            # forward: L + 1 passes reading and writing psi (16 N B each); reverse:
            # L + 1 passes over psi and lambda (32 N B); per circuit 48 N B more
            # (the start state, lambda = H psi)
            byts = (48 * L + 96 * len(cls)) * N
            bound, peak = "hbm", HBM_PEAK_TBPS * 1e12
        if bound == "hbm":
            launch_bytes += byts
        ach = byts / (cms * 1e-3)
        classes[str(n)] = {"circuits": len(cls), "layers": L, "ms": round(cms, 3), "tile_bits": b,
                           "block_threads": ct["block_threads"], "group_width": ct["group_width"],
                           "circuits_per_s": round(len(cls) / (cms * 1e-3), 1), "bound": bound,
                           "achieved_TBps": round(ach / 1e12, 3), "frac": round(ach / peak, 3)}
    out["size_classes"] = classes
    # the whole launch: HBM bytes of its tiled circuits over the launch time
    out["launch_hbm_TBps"] = round(launch_bytes / (ms * 1e-3) / 1e12, 3)
    out["launch_hbm_frac"] = round(launch_bytes / (ms * 1e-3) / (HBM_PEAK_TBPS * 1e12), 3)
    # The launch's roofline: every class at its own bound (HBM for circuits
    # wider than the tile, LDS for those resident in it); the floor is the sum
    # of each class's bytes over its peak, and frac = floor / launch time.
    # achieved is the floor time's HBM-equivalent bytes over the launch time.
    floor_s = sum(c["achieved_TBps"] * 1e12 * c["ms"] * 1e-3 / (lds_peak if c["bound"] == "lds" else
                                                              HBM_PEAK_TBPS * 1e12) for c in classes.values())
    t_launch = ms * 1e-3
    out["roofline"] = {"bound": "hbm", "kernel": "k_batch_hea (one launch, 4096 circuits)",
                       "achieved": round(floor_s * HBM_PEAK_TBPS * 1e12 / t_launch / 1e9, 1),
                       "peak": HBM_PEAK_TBPS * 1000, "unit": "GB/s", "frac": round(floor_s / t_launch, 4),
                       "traffic": None, "avg_ms": round(ms, 3), "launches_timed": 1,
                       "note": "classes wider than the tile at the HBM roof, resident ones at the LDS roof "
                               "(256 CUs x 128 B/clk x 2.4 GHz); bytes per class in size_classes"}
    out["error"] = {"what": "energies and gradients of 12 circuits (stratified) vs the single-circuit complex128 "
                            "adjoint", "max_rel_l2_grad": out["max_rel_grad_err_vs_complex128"],
                    "bar": "relative L2 <= 1e-4"}
    if cpu:
        from oracle import oracle as orc
        core = orc.oracle()
        sample = [specs[i] for i in range(count) if specs[i][0] in (10, 12)][:4]
        t0 = time.perf_counter()
        cnt = 0
        for n, l, p in sample:
            orc.grad_parameter_shift(core, n, lambda x, n=n, l=l: orc.hea_ops(n, l, x), orc.tfim_terms(n), p,
                                     which=list(range(8)))
            cnt += 8
        t = time.perf_counter() - t0
        out["cpu_baseline"] = cpu_leg(round(cnt / t, 2), "gradients/s", "port",
                                      f"oracle FP64 parameter shift, 8 parameters of each of {len(sample)} circuits "
                                      f"with n in (10, 12) ({t:.2f} s)")
    return out


def batch_class_workload(n: int):
    """One size class of the 4096-circuit batch launched alone with a fixed
    configuration (the widest tile, 512 threads, the widest fusion width): a
    short run for hardware-counter passes."""
    from invariantq_amd import batch, workloads
    specs = [s for s in workloads.batch_specs(4096, seed=3) if s[0] == n]
    if not specs:
        raise SystemExit(f"the batch has no circuit of {n} qubits")
    runner = batch.BatchRunner("complex64")
    cfg = dict(tile_bits={n: runner.max_tile_bits}, block_threads=512, group_width=runner.max_group_width)
    runner.run(specs, solo=False, **cfg)
    _, _, ms = runner.run(specs, solo=False, **cfg)
    L = sum(s[1] for s in specs)
    return {"workload": f"batch size class n={n} alone", "circuits": len(specs), "layers": L, "ms": round(ms, 3),
            "config": {"tile_bits": runner.max_tile_bits, "block_threads": 512, "group_width": runner.max_group_width}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--blocks", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-blocks", type=int, default=2)
    ap.add_argument("--no-c64", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--shift-params", type=int, default=24)
    ap.add_argument("--no-workloads", action="store_true",
                    help="skip the circuit20, dm14 and batch sub-objects of the default line")
    ap.add_argument("--only-fused", action="store_true", help="skip every leg but the complex128 fused pass")
    ap.add_argument("--batch-class", type=int, default=0,
                    help="with --workload batch4096: launch only the circuits of this qubit count, fixed configuration")
    ap.add_argument("--workload", default="fused",
                    choices=["fused", "circuit20", "dm14", "batch4096", "adjoint24", "bandwidth", "primitives"],
                    help="fused: the headline line; circuit20 / dm14 / batch4096 / adjoint24: the other "
                         "workloads, each printed as one JSON line")
    a = ap.parse_args()
    if a.only_fused:
        a.no_c64 = a.no_adjoint = a.no_cpu = a.no_workloads = True
    if a.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU: the statevector is not sharded across GPUs, "
                                   "so only --gpus 1 is measured", "n_gpus": a.gpus}))
        sys.exit(2)

    # The binding first: it loads the HIP runtime the library was built with
    # (ROCm 7.2), and torch then binds to that same copy instead of its own.
    from invariantq_amd.core import binding
    import torch
    if not binding.HAS_NATIVE_CORE:
        raise SystemExit(f"HIP engine not available: {binding.native_core_status()}")
    torch.cuda.init()
    if a.workload == "circuit20":
        print(json.dumps(circuit_workload(binding, torch, a.steps, a.warmup, not a.no_cpu)), flush=True)
        return
    if a.workload == "batch4096":
        if a.batch_class:
            print(json.dumps(batch_class_workload(a.batch_class)), flush=True)
            return
        print(json.dumps(batch_workload(torch, 4096, not a.no_cpu)), flush=True)
        return
    if a.workload == "dm14":
        print(json.dumps(density_matrix_workload(binding, torch, not a.no_cpu)), flush=True)
        return
    if a.workload == "adjoint24":
        print(json.dumps(adjoint_vs_shift(torch, 24, 12, max(10, a.steps), a.shift_params, not a.no_cpu)), flush=True)
        return
    if a.workload == "bandwidth":
        print(json.dumps(bandwidth_workload(binding, torch)), flush=True)
        return
    if a.workload == "primitives":
        print(json.dumps(primitives_workload(binding, torch)), flush=True)
        return
    n = a.qubits
    N = 1 << n

    log(f"fused pass c128: n={n}, blocks={a.blocks}, warmup={a.warmup}, steps={a.steps}")
    wall, per_kernel, targets, norm = time_fused(binding, torch, n, "complex128", a.steps, a.warmup, a.blocks)
    bytes_block = 2 * N * 16
    flops_block = 8 * 32 * N
    value = bytes_block * a.blocks / wall / 1e12
    mean_k = np.array([np.mean(x) for x in per_kernel])
    low = np.array([q[-1] < 10 for q in targets])
    # The launcher streams (non-temporal loads/stores, two batches per
    # iteration) unless three or more targets sit inside one 128-byte line
    # (qubits 0-2 for complex128: four batches, no streaming); the two cases are
    # two instantiations of the kernel, listed apart by rocprofv3.
    streaming = np.array([sum(t < 3 for t in q) <= 2 for q in targets])
    kinds = {"k_gate_mfma<double, 2, false, true>": streaming,
             "k_gate_mfma<double, 4, false, false>": ~streaming,
             "k_gate_mfma, blocks on qubits 0-9": low, "k_gate_mfma, unrestricted blocks": ~low}
    kstats = {}
    for name, sel in kinds.items():
        if sel.any():
            d = float(np.mean(mean_k[sel]))
            kstats[name] = {"launches_per_step": int(sel.sum()), "avg_ms": round(d * 1e3, 4),
                            "achieved_TBps": round(bytes_block / d / 1e12, 4)}
    dom = max(list(kstats)[:2], key=lambda k: kstats[k]["launches_per_step"] * kstats[k]["avg_ms"])
    dom_d = kstats[dom]["avg_ms"] * 1e-3
    log(f"c128 wall {wall * 1e3:.1f} ms/step, {value:.3f} TB/s, norm {norm:.15f}, kernels {kstats}")

    c64 = None
    if not a.no_c64:
        log("fused pass c64")
        wall64, pk64, _, norm64 = time_fused(binding, torch, n, "complex64", a.steps, a.warmup, a.blocks)
        b64 = 2 * N * 8
        mean64 = np.array([np.mean(x) for x in pk64])
        c64 = {"TBps": round(b64 * a.blocks / wall64 / 1e12, 4),
               "pct_hbm_peak": round(100 * b64 * a.blocks / wall64 / 1e12 / HBM_PEAK_TBPS, 2),
               "ms_per_step": round(wall64 * 1e3, 3),
               "fp32_TFLOPs": round(flops_block * a.blocks / wall64 / 1e12, 3),
               # of the FP32 matrix peak, counting the dense complex product
               # (8 flops per complex multiply-add; the kernel does 6 of them)
               "fp32_pct_peak": round(100 * flops_block * a.blocks / wall64 / 1e12 / 157.3, 1),
               "avg_kernel_ms": round(float(np.mean(mean64)) * 1e3, 4), "norm": norm64,
               "roofline": roofline("k_gate_mfma<float, ...> (all 200 blocks)", b64, float(np.mean(mean64)),
                                    int(mean64.size))}
        try:
            c64["error"] = c64_fidelity(binding, torch, n, a.blocks)
        except Exception as exc:  # the statement is reported, never required
            c64["error"] = {"unavailable": str(exc)}
        log(f"c64 {c64}")

    adj = None
    if not a.no_adjoint:
        log("adjoint vs parameter shift: 24 qubits, 12 layers")
        # ten timed gradients (about 70 ms): a steady mean whatever --steps is
        adj = adjoint_vs_shift(torch, 24, 12, 10, a.shift_params, not a.no_cpu)
        log(f"adjoint {adj}")

    cpu, err = None, None
    if not a.no_cpu:
        log(f"cpu baseline and oracle error: {a.cpu_blocks} blocks at n={n}")
        try:
            cpu, err = cpu_fused_baseline(binding, n, a.cpu_blocks)
        except Exception as exc:  # the baseline is reported, never required
            cpu = {"error": str(exc)}
        log(f"cpu {cpu} error {err}")

    subs = {}
    if not a.no_workloads:
        wsteps = max(1, min(a.steps, 10))
        for name, fn in (("circuit20", lambda: circuit_workload(binding, torch, wsteps, 2, not a.no_cpu)),
                         ("dm14", lambda: density_matrix_workload(binding, torch, not a.no_cpu)),
                         ("batch", lambda: batch_workload(torch, 4096, not a.no_cpu))):
            log(f"workload {name}")
            t0 = time.perf_counter()
            try:
                subs[name] = fn()
            except Exception as exc:  # one workload's failure does not hide the others
                subs[name] = {"error": f"{type(exc).__name__}: {exc}"}
            subs[name]["bench_s"] = round(time.perf_counter() - t0, 1)
            log(f"{name} {subs[name]}")

    out = {
        "metric": "30q fused-gate HBM TB/s (% peak); adjoint grads/s vs param-shift, same GPU",
        "value": round(value, 4),
        "unit": "TB/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"fused 5-qubit Haar blocks, {n} qubits, complex128 "
                               f"(even blocks on qubits 0-9, odd unrestricted)",
                   "qubits": n, "blocks_per_step": a.blocks, "precision": "complex128", "parallelism": "single GPU"},
        "pct_hbm_peak": round(100 * value / HBM_PEAK_TBPS, 2),
        "fp64_TFLOPs": round(flops_block * a.blocks / wall / 1e12, 3),
        # of the FP64 matrix peak (78.6 TFLOP/s), counting the dense complex
        # product (8 flops per complex multiply-add; the kernel's 3-multiply
        # form does 6)
        "fp64_pct_peak": round(100 * flops_block * a.blocks / wall / 1e12 / 78.6, 1),
        "gate_applications_per_s": round(a.blocks / wall, 2),
        "roofline": {"bound": "hbm", "achieved": round(bytes_block / dom_d / 1e9, 1), "peak": HBM_PEAK_TBPS * 1000,
                     "unit": "GB/s", "frac": round(bytes_block / dom_d / 1e12 / HBM_PEAK_TBPS, 4),
                     "traffic": measured_traffic("fused_gate_30q") if n == 30 else None, "kernel": dom,
                     "algorithmic_bytes_per_launch": bytes_block},
        "kernels": kstats,
        "error": err,
        "complex64": c64,
        "adjoint24": adj,
        "cpu_baseline": cpu,
    }
    out.update(subs)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
