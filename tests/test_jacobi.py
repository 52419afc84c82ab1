"""invariantq_amd/csrc/jacobi.hpp (the entropy eigenvalue step) against the
oracle's full-scan restatement of the reference's capped Jacobi, bit for bit
(tests/jacobi_check.cpp, built with g++ on the CPU).  The checker is compiled
with the oracle's own flags (-O3 -march=x86-64-v3): both sides then contract
the same products into FMAs, so any difference is the algorithm's."""

import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def test_row_max_jacobi_matches_full_scan(tmp_path):
    from oracle import oracle as orc
    if not orc.ORACLE_LIB.exists():
        orc.build(ref=False)
    exe = tmp_path / "jacobi_check"
    subprocess.run(["g++", "-O3", "-march=x86-64-v3", "-std=c++17", f"-I{ROOT / 'invariantq_amd' / 'csrc'}",
                    "-pthread", str(ROOT / "tests" / "jacobi_check.cpp"), str(orc.ORACLE_LIB), "-o", str(exe),
                    f"-Wl,-rpath,{orc.ORACLE_LIB.parent}"], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr
