"""The batch autotuner's choice logic and cache on the CPU, with a stand-in
runner whose timings are scripted (no GPU): per-n tile widths, the thread
count, the fusion-width stage, the cache round trip, merging a later tuning of
other qubit counts into the entry, and the error when no tile width fits."""

import json

import numpy as np
import pytest

from invariantq_amd import batch


class ScriptedRunner:
    """Times: tile width 12 fastest for n <= 14, 13 above; 512 threads faster;
    fusion width 3 fastest."""
    precision = "complex64"
    max_tile_bits = 13
    max_group_width = 4

    def __init__(self):
        self.calls = 0

    def run(self, specs, hx=0.5, tile_bits=None, block_threads=0, num_blocks=0, group_width=0, solo=True):
        assert not solo  # the autotuner times the launch alone
        self.calls += 1
        ms = 0.0
        for n, l, _ in specs:
            w = (tile_bits or {}).get(n, self.max_tile_bits)
            ms += l * 2.0 ** n * 1e-6 * (1.0 if w == (12 if n <= 14 else 13) else 1.3)
        ms *= 0.9 if block_threads == 512 else 1.0
        ms *= {0: 1.0, 2: 1.2, 3: 0.8, 4: 1.0}[group_width]
        return np.zeros(len(specs)), [np.zeros(2 * n * l) for n, l, _ in specs], ms


def specs_for(ns, per=4):
    rng = np.random.default_rng(0)
    return [(n, 2, rng.uniform(0, 6.3, 4 * n)) for n in ns for _ in range(per)]


def test_choices_and_round_trip(tmp_path):
    cache = tmp_path / "tune.json"
    r = ScriptedRunner()
    t = batch.autotune(r, specs_for([10, 14, 18]), cache=cache, per_n=4)
    assert t["tile_bits"] == {10: 12, 14: 12, 18: 13}
    assert t["block_threads"] == 512 and t["group_width"] == 3 and not t["cached"]
    calls = r.calls
    again = batch.autotune(r, specs_for([10, 18]), cache=cache, per_n=4)
    assert again["cached"] and r.calls == calls
    assert again["group_width"] == 3 and again["block_threads"] == 512 and again["tile_bits"][18] == 13


def test_later_tuning_merges_into_the_entry(tmp_path):
    cache = tmp_path / "tune.json"
    r = ScriptedRunner()
    batch.autotune(r, specs_for([10, 14]), cache=cache, per_n=4)
    batch.autotune(r, specs_for([16]), cache=cache, per_n=4)
    (entry,) = json.loads(cache.read_text()).values()
    assert set(entry["tile_bits"]) == {"10", "14", "16"} and entry["group_width"] == 3
    t = batch.autotune(r, specs_for([10, 14, 16]), cache=cache, per_n=4)
    assert t["cached"] and t["tile_bits"] == {10: 12, 14: 12, 16: 13}


def test_no_tile_width_fits():
    with pytest.raises(ValueError, match="no tile width fits"):
        batch.autotune(ScriptedRunner(), specs_for([30]), cache=None, per_n=2)


class MixRunner(ScriptedRunner):
    """Per circuit, width 13 is 5% faster than 12 at either thread count; but a
    256-thread launch whose tiles are all narrower than 13 runs two
    workgroups per compute unit and takes 0.8 of the time: only the
    whole-sample stage sees that."""

    def run(self, specs, hx=0.5, tile_bits=None, block_threads=0, num_blocks=0, group_width=0, solo=True):
        assert not solo  # the autotuner times the launch alone
        self.calls += 1
        widths = [(tile_bits or {}).get(n, self.max_tile_bits) for n, _, _ in specs]
        ms = sum(l * 2.0 ** n * 1e-6 * (0.95 if w == 13 else 1.0) for (n, l, _), w in zip(specs, widths))
        if block_threads == 256 and max(widths) < 13:
            ms *= 0.8
        return np.zeros(len(specs)), [np.zeros(2 * n * l) for n, l, _ in specs], ms


def test_thread_count_chosen_on_the_whole_sample():
    t = batch.autotune(MixRunner(), specs_for([14, 16, 18]), cache=None, per_n=4)
    assert t["block_threads"] == 256 and max(t["tile_bits"].values()) < 13
    assert set(t["timings"]["block_threads"]) == {"256", "512", "256 narrow", "256 uniform"}
