"""CPU check of the gate kernels' address arithmetic (no GPU): compiles
tests/layout_check.cpp against invariantq_amd/csrc/layout.hpp with g++ and
replays every index the direct and tiled kernels compute."""

import itertools
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = tmp_path_factory.mktemp("layout") / "layout_check"
    subprocess.run(["g++", "-O2", "-std=c++17", f"-I{ROOT / 'invariantq_amd' / 'csrc'}",
                    str(ROOT / "tests" / "layout_check.cpp"), "-o", str(exe)], check=True)
    return exe


def _run(exe, cases):
    text = "\n".join(f"{n} {len(q)} " + " ".join(map(str, q)) for n, q in cases)
    r = subprocess.run([str(exe)], input=text, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr


def test_layout_exhaustive_small(checker):
    cases = []
    n = 11
    for K in range(1, 6):
        for q in itertools.permutations(range(n), K):
            if K >= 4 and sum(q) % 7:  # thin the largest sets; orderings still vary
                continue
            cases.append((n, list(q)))
    _run(checker, cases)


def test_layout_random_wide(checker):
    rng = np.random.default_rng(0)
    cases = []
    for n in (9, 10, 12, 14, 16):
        for K in range(1, 6):
            for _ in range(30):
                cases.append((n, [int(v) for v in rng.choice(n, K, replace=False)]))
            cases.append((n, list(range(K))))
            cases.append((n, list(range(n - K, n))))
    _run(checker, cases)
