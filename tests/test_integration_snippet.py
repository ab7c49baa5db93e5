"""The drop-in snippet of INTEGRATION.md compiles against include/gdx_col.h (CPU only: g++ -fsyntax-only).

The snippet's first C++ block is what a maintainer pastes into a model-local override of the reference's
runtime (src/swarm/swarm_runtime.cu:357-496); the runtime's own variables get stand-ins of the types the
reference declares for them."""
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent

RUNTIME_STANDINS = """
#include <cstdlib>
#include <iostream>
#include <stdexcept>
#include <string>
using real = double;
struct swarm;
static void runtime_collision_hooks(int N_P, real total_dust_mass, swarm *dev_particle, int &count_col, real &dt_col,
                                    real &clock_dyn, real clock_sim, const char *output_dir, const char *path_of_frame)
{
    (void)clock_sim; (void)output_dir; (void)path_of_frame;
"""


def _snippet() -> str:
    text = (ROOT/"INTEGRATION.md").read_text()
    return re.search(r"```cpp\n(.*?)```", text, re.S).group(1)


@pytest.mark.parametrize("defines", [[], ["-DCOLLISION_KDTREE"], ["-DCOL_DIAGNOSTICS"],
                                     ["-DCOLLISION_MORTON", "-DCOL_DIAGNOSTICS"]])
def test_integration_snippet_compiles(tmp_path, defines):
    body = _snippet()
    includes = "\n".join(l for l in body.splitlines() if l.startswith("#include"))
    code = "\n".join(l for l in body.splitlines() if not l.startswith("#include"))
    src = tmp_path/"snippet.cpp"
    src.write_text(includes + "\n" + RUNTIME_STANDINS + code + "\n    (void)evolve_collisions; (void)invalidate_col_geometry;\n}\n")
    out = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-Wno-unused-variable",
                          f"-I{ROOT/'include'}", *defines, str(src)], capture_output=True, text=True)
    assert out.returncode == 0, out.stderr
