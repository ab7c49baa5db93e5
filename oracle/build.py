"""Build the CPU oracle (test infrastructure only): one shared library per model directory.

    python oracle/build.py [model ...]

g++ -ffp-contract=off keeps every expression in source order; OpenMP parallelises owners and queries
for the CPU baseline. Outputs go to oracle/lib/ (git-ignored).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
MOD = ROOT / "gamedev_amd" / "mod"
LIB = HERE / "lib"


def models() -> list[str]:
    return sorted(p.name for p in MOD.iterdir() if (p / "const_defs.h").exists())


def lib_path(model: str) -> Path:
    return LIB / f"liboracle_{model}.so"


def _one(model: str) -> Path:
    LIB.mkdir(parents=True, exist_ok=True)
    out = lib_path(model)
    tmp = out.with_suffix(".so.tmp")
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-Wall",
           f"-I{MOD / model}", str(HERE / "oracle.cpp"), "-o", str(tmp)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"oracle build failed for {model}:\n{proc.stdout}{proc.stderr}")
    os.replace(tmp, out)
    return out


def build_planted(model: str, plant: int, out_dir: Path) -> Path:
    """The oracle of `model` with planted mistake number `plant` (see ORC_PLANT in oracle.cpp), for the
    tests that show the pins reject it; -O1 and no OpenMP keep the build to a few seconds."""
    out_dir.mkdir(parents=True, exist_ok=True)
    out = out_dir / f"liboracle_{model}_plant{plant}.so"
    cmd = ["g++", "-O1", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", f"-DORC_PLANT={plant}",
           "-Wno-unknown-pragmas", f"-I{MOD / model}", str(HERE / "oracle.cpp"), "-o", str(out)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"planted oracle build failed for {model}/{plant}:\n{proc.stdout}{proc.stderr}")
    return out


def build(targets: list[str] | None = None) -> list[Path]:
    targets = targets or models()
    with cf.ThreadPoolExecutor(min(8, len(targets))) as ex:
        return list(ex.map(_one, targets))


if __name__ == "__main__":
    for p in build(sys.argv[1:] or None):
        print(p)
