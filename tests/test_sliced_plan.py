"""CPU check of the column-sliced CSR plan: compiles tests/sliced_plan_check.cpp
against pdal_amd/csrc/common/sliced_plan.hpp with g++ and runs it (the index
arithmetic of the device routine replayed thread by thread on random
matrices: every position in range, each nonzero exactly once, row sums in the
plan's order, and the refusal of a row with more than 255 nonzeros in one
slice)."""
from __future__ import annotations

import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_sliced_plan_index_arithmetic(tmp_path):
    exe = tmp_path / "sliced_plan_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Werror", f"-I{ROOT / 'pdal_amd' / 'csrc' / 'common'}",
                    str(ROOT / "tests" / "sliced_plan_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=False, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("OK ") and int(out.stdout.split()[1]) >= 11, out.stdout
