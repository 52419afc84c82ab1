"""invariantq_amd/csrc/split.hpp (the split form of the adjoint's tile passes:
each run as g diag(1, p) R diag(1, q), diagonals gathered into per-pass phase
tables) against every run's full matrix, on random passes, with the groups
applied the way the tile kernels apply them (tests/split_check.cpp, built with
g++ on the CPU)."""

import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def test_split_passes_match_full_matrices(tmp_path):
    exe = tmp_path / "split_check"
    subprocess.run(["g++", "-O2", "-std=c++17", f"-I{ROOT / 'invariantq_amd' / 'csrc'}",
                    str(ROOT / "tests" / "split_check.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr
    _, passes, split = r.stdout.split()
    assert int(passes) == 4000 and int(split) > 1000
