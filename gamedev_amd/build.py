"""Build the collision libraries: one shared library per compile-time model directory.

    python -m gamedev_amd.build [model ...]

Each model directory gamedev_amd/mod/<model>/const_defs.h is compiled with hipcc for gfx950 into
gamedev_amd/lib/libgdx_col_<model>.so (in-tree, so it travels with the source tree).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
MOD = PKG / "mod"
LIB = PKG / "lib"
BUILD = PKG / "_build"
SOURCES = ["geometry.hip", "kdtree.hip", "operator.hip", "majorant.hip", "capi.cpp"]
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
ARCH = "gfx950"
COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-parameter",
]


def models() -> list[str]:
    return sorted(p.name for p in MOD.iterdir() if (p / "const_defs.h").exists())


def lib_path(model: str) -> Path:
    return LIB / f"libgdx_col_{model}.so"


def check_path(model: str) -> Path:
    return LIB / f"pair_forms_check_{model}"


def build_check(model: str) -> Path:
    """Host-only build of the pair-form check program of one model (no GPU code, no HIP runtime use)."""
    LIB.mkdir(parents=True, exist_ok=True)
    out = check_path(model)
    tmp = out.with_suffix(".tmp")
    _run([HIPCC, "-x", "hip", "--cuda-host-only", "-O2", "-std=c++17", "-Wall", "-Wno-unused-function",
          f"-I{MOD / model}", f"-I{CSRC}", str(CSRC / "pair_forms_check.cpp"), "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def topk_check_path() -> Path:
    return LIB / "topk_check"


def build_topk_check() -> Path:
    """Host-only build of the top-K merge check program (model independent)."""
    LIB.mkdir(parents=True, exist_ok=True)
    out = topk_check_path()
    tmp = out.with_suffix(".tmp")
    _run([HIPCC, "-x", "hip", "--cuda-host-only", "-O2", "-std=c++17", "-Wall", f"-I{CSRC}",
          str(CSRC / "topk_check.cpp"), "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (proc.stdout + proc.stderr).splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out, file=sys.stderr)


def _compile(model: str, src: str) -> Path:
    obj_dir = BUILD / model
    obj_dir.mkdir(parents=True, exist_ok=True)
    obj = obj_dir / (src + ".o")
    cmd = [HIPCC, *COMMON, f"-I{MOD / model}", f"-I{CSRC}", "-c", str(CSRC / src), "-o", str(obj)]
    _run(cmd)
    return obj


def build(targets: list[str] | None = None, jobs: int | None = None) -> list[Path]:
    """Compile every source of every requested model; always recompiles."""
    targets = targets or models()
    LIB.mkdir(parents=True, exist_ok=True)
    jobs = jobs or max(1, min(len(targets)*len(SOURCES), os.cpu_count() or 4, 8))
    with cf.ThreadPoolExecutor(jobs) as ex:
        futs = {(m, s): ex.submit(_compile, m, s) for m in targets for s in SOURCES}
        checks = [ex.submit(build_check, m) for m in targets] + [ex.submit(build_topk_check)]
        objs = {k: f.result() for k, f in futs.items()}
        for f in checks:
            f.result()
    out = []
    for m in targets:
        so = lib_path(m)
        tmp = so.with_suffix(".so.tmp")
        _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *[str(objs[(m, s)]) for s in SOURCES],
              "-o", str(tmp)])
        os.replace(tmp, so)
        out.append(so)
    return out


if __name__ == "__main__":
    for p in build(sys.argv[1:] or None):
        print(p)
