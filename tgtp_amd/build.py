"""Builds libtgtp_hip.so (gfx950) and the command-line drop-ins in-tree.

    python -m tgtp_amd.build            # library + CLIs + oracle (oracle only if asked)

Host-only C++ is compiled with g++ (OpenMP via libgomp); HIP sources with
hipcc for gfx950 with FP contraction off (the faithful kernels evaluate the
reference's expressions exactly as written) and packed-FP32 instruction
selection disabled.
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
INCLUDE = ROOT / "include"
BUILD = PKG / "_build"
LIB = PKG / "libtgtp_hip.so"
BIN = PKG / "bin"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TGTP_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-std=c++20", "-O3", f"--offload-arch={ARCH}", "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{INCLUDE}", f"-I{CSRC}",
]
HOST_FLAGS = ["-std=c++20", "-O2", "-fPIC", "-fopenmp", "-ffp-contract=off",
              f"-I{INCLUDE}", f"-I{CSRC}"]

HIP_SOURCES = ["capi.hip", "saba_walk.hip", "saba_part.hip", "mt_walk.hip", "tgtp.hip", "tgtp_fast.hip", "fast_block.hip", "taus_fast.hip", "tgtp_batch.hip"]
HOST_SOURCES = ["graph_io.cpp", "host_tgtp.cpp"]
# binary -> (source, extra defines)
CLI_SOURCES = {"rw_experiment_saba": ("cli_rw.cpp", []),
               "rw_experiment": ("cli_rw.cpp", ["-DTGTP_RW_VARIANT=1"]),
               "rw_experiment_simd": ("cli_rw.cpp", ["-DTGTP_RW_VARIANT=2"]),
               "tgtp": ("cli_tgtp.cpp", []), "benchmark": ("cli_benchmark.cpp", []),
               "benchmark_baseline": ("cli_benchmark.cpp", ["-DTGTP_BENCH_BASELINE=1"]),
               "aesc_example": ("aesc_example.cpp", [])}


# stand-alone HIP measurement tools: binary -> source
GPU_TOOLS = {"gather_roof": "gather_roof.hip"}


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + proc.stdout + proc.stderr)
        raise RuntimeError(f"build step failed: {cmd[-1]}")


def _newer(target: Path, deps: list[Path]) -> bool:
    if not target.exists():
        return False
    t = target.stat().st_mtime
    return all(d.stat().st_mtime <= t for d in deps)


def build(force: bool = True, verbose: bool = False) -> Path:
    """Compile every source; with force=False skip up-to-date objects."""
    BUILD.mkdir(exist_ok=True)
    BIN.mkdir(exist_ok=True)
    headers = list(INCLUDE.glob("*.h")) + list(CSRC.glob("*.hpp"))
    jobs = []
    for src in HIP_SOURCES:
        obj = BUILD / (src + ".o")
        if force or not _newer(obj, [CSRC / src] + headers):
            jobs.append([HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    for src in HOST_SOURCES:
        obj = BUILD / (src + ".o")
        if force or not _newer(obj, [CSRC / src] + headers):
            jobs.append(["g++", *HOST_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        for f in [ex.submit(_run, j) for j in jobs]:
            f.result()
    objs = [str(BUILD / (s + ".o")) for s in HIP_SOURCES + HOST_SOURCES]
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp), *objs, "-lgomp"])
    os.replace(tmp, LIB)
    clis = []
    for name, (src, defines) in CLI_SOURCES.items():
        if not (CSRC / src).exists():
            continue
        clis.append(["g++", "-std=c++20", "-O2", *defines, f"-I{INCLUDE}", str(CSRC / src), "-o",
                     str(BIN / name), f"-L{PKG}", "-ltgtp_hip", f"-Wl,-rpath,{PKG}",
                     "-Wl,-rpath,$ORIGIN/..", "-fopenmp"])
    # measurement tools (HIP, stand-alone)
    for name, src in GPU_TOOLS.items():
        clis.append([HIPCC, *HIP_FLAGS, str(CSRC / src), "-o", str(BIN / name)])
    with cf.ThreadPoolExecutor(max_workers=4) as ex:
        for f in [ex.submit(_run, j) for j in clis]:
            f.result()
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_oracle() -> None:
    """The CPU restatement (test infrastructure) and, when the reference tree
    is present, the reference's own code under oracle/_ref."""
    make = shutil.which("make") or "make"
    _run([make, "-s", "-C", str(ROOT / "oracle"), "-B", "liboracle.so"])
    if Path("/root/reference/src/aesc_baseline.h").exists():
        _run([make, "-s", "-C", str(ROOT / "oracle"), "-B", "ref"])


if __name__ == "__main__":
    build(force="--incremental" not in sys.argv, verbose=True)
    if "--oracle" in sys.argv:
        build_oracle()
