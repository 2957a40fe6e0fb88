"""Builds the native pieces in-tree.

- ``poissonrecon_gpu_amd/libprg.so``: the HIP library (all ``csrc`` sources),
  compiled with hipcc for gfx950;
- ``poissonrecon_gpu_amd/poisson_recon``: the command-line front end;
- ``oracle/liboracle.so``: the CPU oracle used by the tests (g++).

Usage: ``python -m poissonrecon_gpu_amd.build`` (or ``build_all()``).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = ROOT / "_build"
LIB = PKG / "libprg.so"
CLI = PKG / "poisson_recon"
ORACLE_DIR = ROOT / "oracle"
ORACLE_LIB = ORACLE_DIR / "liboracle.so"

ARCH = os.environ.get("PRG_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    # bit-exact float sequences: no contraction into FMA anywhere
    "-ffp-contract=off",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-parameter",
    f"-I{ROOT / 'include'}",
]
DEVICE = [
    f"--offload-arch={ARCH}",
    # no packed FP32 instructions (v_pk_*_f32) in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")


def _compile(src: Path) -> Path:
    BUILD.mkdir(exist_ok=True)
    obj = BUILD / (src.name + ".o")
    cmd = [HIPCC, *COMMON, "-x", "hip", *DEVICE, "-c", str(src), "-o", str(obj)]
    _run(cmd)
    return obj


def build_library(jobs: int = 8) -> Path:
    srcs = sorted(p for p in CSRC.iterdir() if p.suffix in (".hip", ".cpp") and p.name != "cli.cpp")
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, *DEVICE, "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


def build_cli() -> Path:
    src = CSRC / "cli.cpp"
    if not src.exists():
        return CLI
    _run([HIPCC, *COMMON, str(src), "-o", str(CLI), f"-L{PKG}", "-lprg", "-Wl,-rpath,$ORIGIN"])
    return CLI


def build_oracle() -> Path:
    _run(["make", "-s", "-C", str(ORACLE_DIR), "-B", "liboracle.so"])
    return ORACLE_LIB


def build_all(jobs: int = 8) -> None:
    build_oracle()
    build_library(jobs)
    build_cli()


if __name__ == "__main__":
    build_all(int(sys.argv[1]) if len(sys.argv) > 1 else 8)
    print(f"built {LIB}, {CLI}, {ORACLE_LIB}")
