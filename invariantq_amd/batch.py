"""Batched variational sweep in one persistent launch, with an autotuner.

``BatchRunner.run`` sends thousands of ansatz circuits (RY, RZ on every qubit
and a CZ ring per layer; open-chain TFIM observable) of uneven qubit count and
depth to the GPU kernel ``k_batch_hea``: energies and complete adjoint
gradients come back in one launch.  ``autotune`` chooses the LDS tile width
per qubit count and the workgroup size by timing candidates on a sample of
the batch, and caches the choice per (device, precision) in a JSON file.

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

    def __del__(self):
        if getattr(self, "_ptr", None):
            self._free(self._ptr)
            self._ptr = None

    def run(self, specs: Sequence[Tuple[int, int, np.ndarray]], hx: float = 0.5,
            tile_bits: Optional[Dict[int, int]] = None, block_threads: int = 0, num_blocks: int = 0):
        """specs: (n, layers, params) per circuit.  Returns (energies, [grad per circuit], kernel ms)."""
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
    try:
        import torch
        name = torch.cuda.get_device_properties(0).gcnArchName
    except Exception:  # the cache key only needs to be stable per machine
        name = "gfx950"
    return f"{name}:{precision}"


def autotune(runner: BatchRunner, specs: Sequence[Tuple[int, int, np.ndarray]], hx: float = 0.5,
             cache: Optional[Path] = DEFAULT_CACHE, per_n: int = 256, force: bool = False) -> Dict:
    """Choose the tile width for every qubit count present in `specs` and the
    workgroup size, by timing candidates on up to `per_n` circuits of each count.

    Candidates: tile width in {max, max-1, max-2} (circuits wider than the tile
    take the two-pass tiled path; a width >= n keeps the circuit in LDS) and
    256 or 512 threads per workgroup.  Returns {"tile_bits": {n: width},
    "block_threads": t, "timings": ...} and stores it in `cache`.
    """
    key = _device_key(runner.precision)
    table = {}
    if cache is not None and Path(cache).exists():
        try:
            table = json.loads(Path(cache).read_text())
        except (OSError, ValueError):
            table = {}
    ns = sorted({int(s[0]) for s in specs})
    entry = table.get(key)
    if entry and not force and all(str(n) in entry["tile_bits"] for n in ns):
        return {"tile_bits": {int(k): v for k, v in entry["tile_bits"].items()},
                "block_threads": entry["block_threads"], "timings": entry.get("timings", {}), "cached": True}
    bmax = runner.max_tile_bits
    timings: Dict[str, Dict[str, float]] = {}
    best: Dict[int, int] = {}
    per_threads = {256: 0.0, 512: 0.0}
    for n in ns:
        sample = [s for s in specs if int(s[0]) == n][:per_n]
        cands = sorted({w for w in (bmax, bmax - 1, bmax - 2) if 2 * w >= n})
        res = {}
        for w in cands:
            for t in (256, 512):
                runner.run(sample[:8], hx, {n: w}, t)  # warm-up
                _, _, ms = runner.run(sample, hx, {n: w}, t)
                res[f"{w}/{t}"] = ms
        timings[str(n)] = res
        w_best = min(cands, key=lambda w: min(res[f"{w}/256"], res[f"{w}/512"]))
        best[n] = w_best
        for t in per_threads:
            per_threads[t] += res[f"{w_best}/{t}"]
    threads = min(per_threads, key=per_threads.get)
    out = {"tile_bits": best, "block_threads": threads, "timings": timings, "cached": False}
    if cache is not None:
        try:
            Path(cache).parent.mkdir(parents=True, exist_ok=True)
            table[key] = {"tile_bits": {str(k): v for k, v in best.items()}, "block_threads": threads,
                          "timings": timings}
            Path(cache).write_text(json.dumps(table, indent=1))
        except OSError:
            pass
    return out
