"""Batched variational sweep in one persistent launch, with an autotuner.

``BatchRunner.run`` sends thousands of ansatz circuits (RY, RZ on every qubit
and a CZ ring per layer; open-chain TFIM observable) of uneven qubit count and
depth to the GPU kernel ``k_batch_hea``: energies and complete adjoint
gradients come back in one launch.  ``autotune`` chooses, by timing
candidates on a sample of the batch, the LDS tile width per qubit count
(amplitudes per workgroup), the workgroup size, and the fusion width (qubits
whose rotations are fused into one register pass of the tile sweeps), and
caches the choice per (device, precision) in a JSON file.

No reference counterpart: the reference evaluates one circuit per simulator
call (invariantq/benchmarks/vqe.py:44-62).
"""

from __future__ import annotations

import ctypes
import json
import os
from pathlib import Path
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .core import binding

_d = ctypes.POINTER(ctypes.c_double)
_i32 = ctypes.POINTER(ctypes.c_int32)

DEFAULT_CACHE = Path(os.environ.get("IQ_AUTOTUNE_CACHE", Path.home() / ".cache" / "invariantq_amd" / "autotune.json"))


def _dp(a):
    return a.ctypes.data_as(_d)


class BatchRunner:
    def __init__(self, precision: str = "complex64"):
        L = binding.lib()
        self.precision = precision
        self.prec_code = binding.PRECISIONS[precision]
        L.iq_clear_error()
        self._free = L.iq_batch_free  # held here: module globals may be gone when __del__ runs at exit
        self._ptr = L.iq_batch_create(self.prec_code)
        if not self._ptr:
            binding.check_error()
            raise MemoryError("could not create the batch context")
        self.max_tile_bits = L.iq_batch_max_tile_bits(self.prec_code)
        self.max_group_width = L.iq_batch_max_group_width(self.prec_code)

    def __del__(self):
        if getattr(self, "_ptr", None):
            self._free(self._ptr)
            self._ptr = None

    # One circuit per workgroup: a circuit runs on one compute unit, so a
    # size class with few circuits leaves the GPU idle for its whole sweep
    # (one 20-qubit, 6-layer circuit: 36 ms, where the single-circuit engine,
    # every compute unit on one circuit, takes about 0.5 ms).  Per circuit on
    # one CU (complex64, measured at n = 14, 17, 20): about 36 ms * 2^(n-20)
    # * layers / 6; on the single-circuit engine about 0.28 ms + 3.5e-8 ms *
    # 2^n * layers, plus making the circuit for a shape not already kept.
    _CU_MS_20_6 = 36.0
    _ENGINE_MS_FIXED, _ENGINE_MS_PER_AMP_LAYER = 0.28, 3.5e-8
    # making one such circuit (its tape, observable and first plan): about
    # 5 ms at n = 20, 6 layers
    _BUILD_MS_FIXED, _BUILD_MS_PER_GATE = 2.0, 0.025
    _SOLO_CACHE = 8  # circuits kept for later calls (least recently used dropped)

    def solo_classes(self, specs: Sequence[Tuple[int, int, np.ndarray]]) -> List[int]:
        """Qubit counts whose circuits the single-circuit engine evaluates
        in less than half the time their persistent-launch sweep would hold
        a compute unit (a class with fewer circuits than fill the GPU)."""
        by_n: Dict[int, List[int]] = {}
        for n, layers, _ in specs:
            by_n.setdefault(int(n), []).append(int(layers))
        cached = self.__dict__.get("_solo", {})
        out = []
        for n, ls in by_n.items():
            if n < 16:
                continue
            one_cu = self._CU_MS_20_6 * 2.0 ** (n - 20) * max(ls) / 6.0
            engine = sum(self._ENGINE_MS_FIXED + self._ENGINE_MS_PER_AMP_LAYER * 2.0 ** n * l for l in ls)
            engine += sum(self._BUILD_MS_FIXED + self._BUILD_MS_PER_GATE * 2 * n * l
                          for l in set(ls) if not any(k[0] == n and k[1] == l for k in cached))
            if engine < 0.5 * one_cu:
                out.append(n)
        return sorted(out)

    def _solo_circuit(self, n: int, layers: int, hx: float):
        """The single-circuit engine's circuit for one (n, layers, hx).  Each
        holds device memory (its state and the stored states of its sweep), so
        only the _SOLO_CACHE most recently used are kept between calls."""
        from collections import OrderedDict
        from .adjoint import GpuCircuit
        from .workloads import Ansatz, tfim_terms
        cache = self.__dict__.setdefault("_solo", OrderedDict())
        key = (n, layers, float(hx))
        if key in cache:
            cache.move_to_end(key)
        else:
            c = GpuCircuit.from_ansatz(Ansatz(n, layers), self.precision)
            c.set_observable(tfim_terms(n, hx))
            cache[key] = c
            while len(cache) > self._SOLO_CACHE:
                cache.popitem(last=False)
        return cache[key]

    def run(self, specs: Sequence[Tuple[int, int, np.ndarray]], hx: float = 0.5,
            tile_bits: Optional[Dict[int, int]] = None, block_threads: int = 0, num_blocks: int = 0,
            group_width: int = 0, solo: bool = True):
        """specs: (n, layers, params) per circuit; group_width 0 = the widest.
        Returns (energies, [grad per circuit], kernel ms).  With `solo`, the
        circuits of the size classes solo_classes() names go through the
        single-circuit adjoint engine (same ansatz, observable, precision and
        parameter layout), the rest through the persistent launch; the kernel
        ms is the launch's."""
        for n, layers, params in specs:
            if np.asarray(params).size != 2 * int(n) * int(layers):
                raise ValueError("each circuit needs 2 * n * layers parameters")
        solo_n = set(self.solo_classes(specs)) if solo else set()
        self.last_solo = sum(1 for sp in specs if int(sp[0]) in solo_n)
        if solo_n:
            energies = np.empty(len(specs), dtype=np.float64)
            grads: List[Optional[np.ndarray]] = [None] * len(specs)
            rest = [i for i, sp in enumerate(specs) if int(sp[0]) not in solo_n]
            ms = 0.0
            if rest:
                e_r, g_r, ms = self._launch([specs[i] for i in rest], hx, tile_bits, block_threads, num_blocks,
                                            group_width)
                for k, i in enumerate(rest):
                    energies[i], grads[i] = e_r[k], g_r[k]
            for i, (n, layers, params) in enumerate(specs):
                if int(n) in solo_n:
                    c = self._solo_circuit(int(n), int(layers), hx)
                    e, g = c.gradient(np.asarray(params, dtype=np.float64))
                    energies[i], grads[i] = e, np.asarray(g, dtype=np.float64)
            return energies, grads, ms
        return self._launch(specs, hx, tile_bits, block_threads, num_blocks, group_width)

    def _launch(self, specs, hx, tile_bits, block_threads, num_blocks, group_width):
        if binding._NATIVE_LIB.iq_batch_set_group_width(self._ptr, int(group_width)) != 0:
            binding.check_error()
        nq = np.ascontiguousarray([s[0] for s in specs], dtype=np.int32)
        ly = np.ascontiguousarray([s[1] for s in specs], dtype=np.int32)
        sizes = 2 * nq.astype(np.int64) * ly
        params = np.ascontiguousarray(np.concatenate([np.asarray(s[2], dtype=np.float64).reshape(-1) for s in specs]))
        if params.size != int(sizes.sum()):
            raise ValueError("each circuit needs 2 * n * layers parameters")
        tb = np.ascontiguousarray([(tile_bits or {}).get(int(n), 0) for n in nq], dtype=np.int32)
        grads = np.empty(params.size, dtype=np.float64)
        energies = np.empty(len(specs), dtype=np.float64)
        ms = ctypes.c_double(0.0)
        binding._NATIVE_LIB.iq_batch_hea_tfim(self._ptr, len(specs), nq.ctypes.data_as(_i32), ly.ctypes.data_as(_i32),
                                              _dp(params), hx, tb.ctypes.data_as(_i32), block_threads, num_blocks,
                                              _dp(energies), _dp(grads), ctypes.byref(ms))
        binding.check_error()
        offs = np.concatenate([[0], np.cumsum(sizes)])
        return energies, [grads[offs[i]:offs[i + 1]] for i in range(len(specs))], ms.value


def _device_key(precision: str) -> str:
    """Cache key: the library's backend (its one target architecture) and the precision."""
    name = binding.lib().iq_backend_name()
    name = name.decode() if isinstance(name, bytes) else str(name)
    return f"{name}:{precision}"


def _load_table(cache: Optional[Path]) -> Dict:
    if cache is None or not Path(cache).exists():
        return {}
    try:
        table = json.loads(Path(cache).read_text())
        return table if isinstance(table, dict) else {}
    except (OSError, ValueError):
        return {}


def autotune(runner: BatchRunner, specs: Sequence[Tuple[int, int, np.ndarray]], hx: float = 0.5,
             cache: Optional[Path] = DEFAULT_CACHE, per_n: int = 256, force: bool = False) -> Dict:
    """Choose the tile width for every qubit count present in `specs`, the
    workgroup size and the fusion width, by timing candidates on up to `per_n`
    circuits of each count.

    Stage 1 (default fusion width): tile width in {max, max-1, max-2} (circuits
    wider than the tile take the two-pass tiled path; a width >= n keeps the
    circuit in LDS; widths below n/2 do not fit) x 256 or 512 threads per
    workgroup, timed per n; for each thread count, the fastest width per n.
    Stage 2: the thread count, each with its widths (and 256 threads with
    every width below the widest, per n or all at the next-widest, so that
    two workgroups share a compute unit), timed on the whole batch (up to 16384 circuits; beyond, on the
    per-n samples together): the launch sizes LDS for its widest tile, and
    the mix of widths and the number of circuits per compute unit, not one
    qubit count, decide how many workgroups share a compute unit.
    Stage 3: with those, the fusion width (qubits per register pass of the
    tile sweeps, 2 up to the precision's widest), timed the same way.
    Returns {"tile_bits": {n: width}, "block_threads": t, "group_width": g,
    "timings": ..., "cached": bool}.  The cache keeps one entry per (device,
    precision): tile widths of qubit counts tuned earlier are kept and new
    ones merged in; the thread count and fusion width are those of the
    latest tuning.
    """
    key = _device_key(runner.precision)
    table = _load_table(cache)
    ns = sorted({int(s[0]) for s in specs})
    entry = table.get(key) or {}
    known = entry.get("tile_bits", {})
    if (entry and not force and all(str(n) in known for n in ns) and "group_width" in entry
            and "block_threads" in entry):
        return {"tile_bits": {int(k): v for k, v in known.items()}, "block_threads": entry["block_threads"],
                "group_width": entry["group_width"], "timings": entry.get("timings", {}), "cached": True}
    bmax = runner.max_tile_bits
    timings: Dict[str, Dict[str, float]] = {}
    best_by_t: Dict[int, Dict[int, int]] = {256: {}, 512: {}, -256: {}}  # -256: 256 threads, narrow tiles
    samples = {}
    for n in ns:
        sample = [s for s in specs if int(s[0]) == n][:per_n]
        samples[n] = sample
        cands = sorted({w for w in (bmax, bmax - 1, bmax - 2) if 2 * w >= n})
        if not cands:
            raise ValueError(f"no tile width fits {n} qubits: the tiled path needs 2 * width >= n, "
                             f"and the widest tile holds {bmax} qubits")
        res = {}
        for w in cands:
            for t in (256, 512):
                runner.run(sample[:8], hx, {n: w}, t, solo=False)  # warm-up
                _, _, ms = runner.run(sample, hx, {n: w}, t, solo=False)
                res[f"{w}/{t}"] = ms
        timings[str(n)] = res
        for t in (256, 512):
            best_by_t[t][n] = min(cands, key=lambda w: res[f"{w}/{t}"])
        # below the widest tile, two 256-thread workgroups fit a compute unit
        narrow = [w for w in cands if w < bmax] or cands
        best_by_t[-256][n] = min(narrow, key=lambda w: res[f"{w}/256"])
    # and every qubit count at the one width below the widest (where it fits)
    best_by_t[-257] = {n: (min(n, bmax - 1) if 2 * (bmax - 1) >= n else best_by_t[256][n]) for n in ns}
    # stages 2 and 3 time the batch itself (its occupancy is what they tune),
    # or for very large batches the per-n samples together
    everything = list(specs) if len(specs) <= 16384 else [s for n in ns for s in samples[n]]
    def steady(widths, threads, group_width=0):
        # the second of two back-to-back launches: a launch right after a small
        # one can run faster than the same launch repeated, as in use
        runner.run(everything, hx, widths, threads, group_width=group_width, solo=False)
        return runner.run(everything, hx, widths, threads, group_width=group_width, solo=False)[2]

    t_ms = {}
    for key_t, widths in best_by_t.items():
        t_ms[key_t] = steady(widths, 256 if key_t < 0 else key_t)
    labels = {256: "256", 512: "512", -256: "256 narrow", -257: "256 uniform"}
    timings["block_threads"] = {labels[t]: ms for t, ms in t_ms.items()}
    choice = min(t_ms, key=t_ms.get)
    threads, best = (256 if choice < 0 else choice), best_by_t[choice]
    gw_ms = {}
    for g in range(2, runner.max_group_width + 1):
        gw_ms[g] = steady(best, threads, g)
    group_width = min(gw_ms, key=gw_ms.get)
    timings["group_width"] = {str(g): ms for g, ms in gw_ms.items()}
    out = {"tile_bits": best, "block_threads": threads, "group_width": group_width, "timings": timings,
           "cached": False}
    if cache is not None:
        try:
            merged = {str(k): v for k, v in known.items()}
            merged.update({str(k): v for k, v in best.items()})
            table[key] = {"tile_bits": merged, "block_threads": threads, "group_width": group_width,
                          "timings": timings}
            Path(cache).parent.mkdir(parents=True, exist_ok=True)
            Path(cache).write_text(json.dumps(table, indent=1))
        except OSError:
            pass
    return out
