"""Build the CPU oracle shared objects (test infrastructure, gcc only).

Output: oracle/_build/libcp_al_oracle.so, libfa_cp_oracle.so,
libfa_cp_oracle_rev.so (the FA_CP oracle summing in reverse order, the
summation-order noise floor of a trajectory) and liblin_cp_al_oracle.so (the
linearized CP-AL QP solver, the FA_CP source built with -DORACLE_LIN_CP_AL) and
libplanted_fa_cp_oracle.so (the planted-KKT FA_CP, -DORACLE_PLANTED_FA_CP) and
libplanted_cp_al_oracle.so (the planted-KKT CP-AL, -DORACLE_PLANTED_CP_AL), each
with a reverse-summing build, libcp_al_oracle_omp.so (the LP oracle at -O3
with OpenMP: the CPU baseline of the benchmark) and cp_al_oracle_cli (the LP
runner's own source linked against the LP oracle).  The oracle exports the same C names
as the product library; it is built with hidden visibility and -Bsymbolic so a
process can load both without the two binding to each other's symbols.

The reference itself cannot be compiled here (CUDA / cuSPARSE / cuBLAS sources,
PSLP fetched over the network), so there is no oracle/_ref build.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
OUT = HERE / "_build"

FLAGS = ["-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fvisibility=hidden",
         "-Wl,-Bsymbolic", f"-I{ROOT / 'include'}"]


def _one(src: str, name: str, defines=()) -> Path:
    OUT.mkdir(exist_ok=True)
    lib = OUT / name
    tmp = OUT / (name + ".tmp")
    cmd = ["g++", *FLAGS, *defines, str(HERE / src), "-o", str(tmp)]
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, lib)
    return lib


def build() -> Path:
    # the CPU baseline's build: -O3 for an x86-64-v3 host and OpenMP over the
    # host's cores (same source, same results: no FMA contraction, no reassociation)
    _one("cp_al_oracle.cpp", "libcp_al_oracle_omp.so", ["-O3", "-march=x86-64-v3", "-fopenmp"])
    _one("fa_cp_oracle.cpp", "libfa_cp_oracle.so")
    _one("fa_cp_oracle.cpp", "libfa_cp_oracle_rev.so", ["-DFA_CP_ORACLE_REVERSE_SUMS"])
    _one("fa_cp_oracle.cpp", "liblin_cp_al_oracle.so", ["-DORACLE_LIN_CP_AL"])
    _one("fa_cp_oracle.cpp", "libplanted_fa_cp_oracle.so", ["-DORACLE_PLANTED_FA_CP"])
    _one("fa_cp_oracle.cpp", "libplanted_fa_cp_oracle_rev.so", ["-DORACLE_PLANTED_FA_CP", "-DFA_CP_ORACLE_REVERSE_SUMS"])
    _one("fa_cp_oracle.cpp", "libplanted_cp_al_oracle.so", ["-DORACLE_PLANTED_CP_AL"])
    _one("fa_cp_oracle.cpp", "libplanted_cp_al_oracle_rev.so", ["-DORACLE_PLANTED_CP_AL", "-DFA_CP_ORACLE_REVERSE_SUMS"])
    lib = _one("cp_al_oracle.cpp", "libcp_al_oracle.so")
    # the LP runner's source (policy layer, MPS reader, output files) linked
    # against the oracle: the runner's routes and phases with oracle solves,
    # the checker of the GPU runner
    csrc = ROOT / "pdal_amd" / "csrc" / "lp"
    cli = OUT / "cp_al_oracle_cli"
    cmd = ["g++", "-O2", "-std=c++17", "-ffp-contract=off", f"-I{ROOT / 'include'}", str(csrc / "cli.cpp"),
           str(csrc / "mps_parser.cpp"), f"-L{OUT}", "-l:libcp_al_oracle.so", "-Wl,-rpath,$ORIGIN", "-lz",
           "-o", str(cli) + ".tmp"]
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(str(cli) + ".tmp", cli)
    return lib


if __name__ == "__main__":
    build()
    sys.exit(0)
