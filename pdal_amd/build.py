"""Build the native libraries of pdal_amd in-tree.

Products (HIP, gfx950):
  pdal_amd/lib/libcp_al.so   -- CP_AL LP solver behind include/cp_al.h
  pdal_amd/lib/cp_al         -- the LP runner command line
  pdal_amd/lib/libfa_cp.so   -- FA_CP QP solver behind include/fa_cp.h (+ the QPS reader, fa_cp_mps_parser.h)
  pdal_amd/lib/fa_cp         -- the QP runner command line
  pdal_amd/lib/libpdal_ws.so -- learned warm start (MLP on MFMA, POD expansion), include/pdal_warmstart.h
Checkers (CPU, gcc), built by oracle/build_oracle.py, are separate.

Run `python -m pdal_amd.build` (or __graft_entry__.build()).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "pdal_amd" / "csrc"
LIB = ROOT / "pdal_amd" / "lib"
INCLUDE = ROOT / "include"

ARCH = os.environ.get("PDAL_OFFLOAD_ARCH", "gfx950")

# Packed-FP32 instructions are kept out of every HIP build.
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fvisibility=hidden",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{INCLUDE}",
]

CP_AL_SOURCES = [
    CSRC / "lp" / "cp_al_api.cpp",
    CSRC / "lp" / "cp_al_solver.hip",
    CSRC / "lp" / "batch_solver.hip",
    CSRC / "lp" / "mps_parser.cpp",
    CSRC / "lp" / "presolve.cpp",
]


def _hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    return exe


def _run(cmd: list[str]) -> None:
    print("+", " ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], check=True)


def _compile_objects(sources, outdir: Path, extra=()) -> list[Path]:
    outdir.mkdir(parents=True, exist_ok=True)
    objs = []
    for src in sources:
        obj = outdir / (src.stem + src.suffix.replace(".", "_") + ".o")
        lang = ["-x", "hip"] if src.suffix == ".hip" else ["-x", "c++"]
        _run([_hipcc(), *HIP_FLAGS, *extra, *lang, "-c", src, "-o", obj])
        objs.append(obj)
    return objs


def build_cp_al() -> Path:
    objdir = ROOT / "pdal_amd" / "_build" / "cp_al"
    objs = _compile_objects(CP_AL_SOURCES, objdir)
    LIB.mkdir(parents=True, exist_ok=True)
    out = LIB / "libcp_al.so"
    tmp = LIB / "libcp_al.so.tmp"
    _run([_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,-Bsymbolic", *objs, "-lz",
          "-o", tmp])
    os.replace(tmp, out)
    # LP runner CLI linked against the shared library
    cli_src = CSRC / "lp" / "cli.cpp"
    if cli_src.exists():
        exe = LIB / "cp_al"
        _run(["g++", "-O2", "-std=c++17", f"-I{INCLUDE}", cli_src, f"-L{LIB}", "-lcp_al",
              f"-Wl,-rpath,$ORIGIN", "-o", exe])
    return out


FA_CP_SOURCES = [
    CSRC / "qp" / "fa_cp_api.cpp",
    CSRC / "qp" / "fa_cp_solver.hip",
    CSRC / "qp" / "qps_parser.cpp",
]


def build_fa_cp() -> Path:
    objdir = ROOT / "pdal_amd" / "_build" / "fa_cp"
    objs = _compile_objects(FA_CP_SOURCES, objdir)
    LIB.mkdir(parents=True, exist_ok=True)
    out = LIB / "libfa_cp.so"
    tmp = LIB / "libfa_cp.so.tmp"
    _run([_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,-Bsymbolic", *objs, "-lz", "-o", tmp])
    os.replace(tmp, out)
    # QP runner CLI linked against the shared library
    _run(["g++", "-O2", "-std=c++17", f"-I{INCLUDE}", CSRC / "qp" / "cli.cpp", f"-L{LIB}", "-lfa_cp",
          f"-Wl,-rpath,$ORIGIN", "-o", LIB / "fa_cp"])
    return out


def build_lin_cp_al() -> Path:
    """The linearized CP-AL QP solver: the FA_CP sources built with -DPDAL_LIN_CP_AL."""
    objdir = ROOT / "pdal_amd" / "_build" / "lin_cp_al"
    objs = _compile_objects(FA_CP_SOURCES, objdir, extra=["-DPDAL_LIN_CP_AL"])
    LIB.mkdir(parents=True, exist_ok=True)
    out = LIB / "librhr_lin_cp_al.so"
    tmp = LIB / "librhr_lin_cp_al.so.tmp"
    _run([_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,-Bsymbolic", *objs, "-lz", "-o", tmp])
    os.replace(tmp, out)
    # its runner (QP/LIN_CP_AL's rhr_lin_cp_al) linked against the shared library
    _run(["g++", "-O2", "-std=c++17", "-DPDAL_LIN_CP_AL", f"-I{INCLUDE}", CSRC / "qp" / "variant_cli.cpp",
          f"-L{LIB}", "-l:librhr_lin_cp_al.so", f"-Wl,-rpath,$ORIGIN", "-o", LIB / "rhr_lin_cp_al"])
    return out


def build_planted_fa_cp() -> Path:
    """The planted-KKT FA_CP (RHR-FA-CP): the FA_CP sources built with -DPDAL_PLANTED_FA_CP."""
    objdir = ROOT / "pdal_amd" / "_build" / "planted_fa_cp"
    objs = _compile_objects(FA_CP_SOURCES, objdir, extra=["-DPDAL_PLANTED_FA_CP"])
    LIB.mkdir(parents=True, exist_ok=True)
    out = LIB / "librhr_fa_cp.so"
    tmp = LIB / "librhr_fa_cp.so.tmp"
    _run([_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,-Bsymbolic", *objs, "-lz", "-o", tmp])
    os.replace(tmp, out)
    # its runner (QP/PLANTED_FA_CP's rhr_fa_cp) linked against the shared library
    _run(["g++", "-O2", "-std=c++17", "-DPDAL_PLANTED_FA_CP", f"-I{INCLUDE}", CSRC / "qp" / "variant_cli.cpp",
          f"-L{LIB}", "-l:librhr_fa_cp.so", f"-Wl,-rpath,$ORIGIN", "-o", LIB / "rhr_fa_cp"])
    return out


def build_planted_cp_al() -> Path:
    """The planted-KKT CP-AL (RHR-CP-AL): the FA_CP sources built with -DPDAL_PLANTED_CP_AL."""
    objdir = ROOT / "pdal_amd" / "_build" / "planted_cp_al"
    objs = _compile_objects(FA_CP_SOURCES, objdir, extra=["-DPDAL_PLANTED_CP_AL"])
    LIB.mkdir(parents=True, exist_ok=True)
    out = LIB / "librhr_cp_al.so"
    tmp = LIB / "librhr_cp_al.so.tmp"
    _run([_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,-Bsymbolic", *objs, "-lz", "-o", tmp])
    os.replace(tmp, out)
    # its runner (QP/PLANTED_CP_AL's rhr_cp_al) linked against the shared library
    _run(["g++", "-O2", "-std=c++17", "-DPDAL_PLANTED_CP_AL", f"-I{INCLUDE}", CSRC / "qp" / "variant_cli.cpp",
          f"-L{LIB}", "-l:librhr_cp_al.so", f"-Wl,-rpath,$ORIGIN", "-o", LIB / "rhr_cp_al"])
    return out


def build_ws() -> Path:
    objdir = ROOT / "pdal_amd" / "_build" / "ws"
    objs = _compile_objects([CSRC / "ws" / "warmstart.hip"], objdir)
    LIB.mkdir(parents=True, exist_ok=True)
    out = LIB / "libpdal_ws.so"
    tmp = LIB / "libpdal_ws.so.tmp"
    _run([_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,-Bsymbolic", *objs, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_all() -> None:
    build_cp_al()
    build_fa_cp()
    build_lin_cp_al()
    build_planted_fa_cp()
    build_planted_cp_al()
    build_ws()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
