"""Builds invariantq_amd/core/libinvariantq.so for gfx950 with hipcc.

    python -m invariantq_amd.build          # rebuild everything
    python -m invariantq_amd.build --incremental

Objects go to invariantq_amd/_build/; the shared library lands next to the
binding (invariantq_amd/core/) so it travels with the source tree.
"""

from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "core" / "libinvariantq.so"
ARCH = os.environ.get("IQ_OFFLOAD_ARCH", "gfx950")

SOURCES = ["gate_kernels.hip", "sv_kernels.hip", "engine.hip", "adjoint.hip", "batch.hip", "other_models.cpp"]

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 VALU forms are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-value", "-Wno-unused-result", "-Wno-pass-failed",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]


def _sources():
    return [CSRC / s for s in SOURCES if (CSRC / s).exists()]


def _stale(src: Path, obj: Path) -> bool:
    if not obj.exists():
        return True
    deps = [src] + list(CSRC.glob("*.hpp")) + list((ROOT / "include").glob("*.h"))
    t = obj.stat().st_mtime
    return any(d.stat().st_mtime > t for d in deps)


def _compile(src: Path, obj: Path) -> None:
    cmd = [HIPCC] + CFLAGS + ["-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed on {src.name}")


def build(incremental: bool = False, jobs: int = 8) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = _sources()
    objs = [BUILD / (s.name + ".o") for s in srcs]
    todo = [(s, o) for s, o in zip(srcs, objs) if not incremental or _stale(s, o)]
    if todo:
        with ThreadPoolExecutor(max_workers=max(1, min(jobs, len(todo)))) as ex:
            list(ex.map(lambda so: _compile(*so), todo))
    if todo or not LIB.exists():
        tmp = LIB.with_suffix(".so.tmp")
        cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp)] + [str(o) for o in objs]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("link of libinvariantq.so failed")
        os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--incremental", action="store_true")
    ap.add_argument("-j", "--jobs", type=int, default=8)
    a = ap.parse_args()
    print(build(incremental=a.incremental, jobs=a.jobs))
