"""Builds tangerine_amd/libtangerine_gpu.so for gfx950 with hipcc.

Usage: python -m tangerine_amd.build [--jobs N]

Every HIP translation unit is compiled for gfx950 only, with packed FP32
instructions disabled and FMA contraction off.  The evaluation kernels are
built twice: with reference bit parity (exact) and with explicit fused
multiply-adds, hardware sqrt / reciprocal and float blends (fast).
"""

from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libtangerine_gpu.so"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
HIPCC = str(ROCM / "bin" / "hipcc")

HIP_COMMON = [
    "--offload-arch=gfx950",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + str(CSRC),
    "-Wno-unused-result",
]

# The interpreter dispatches on wave-uniform record heads.  Left to the
# control-flow structurizer, those uniform switches become chains of flag
# blocks (and the P-points-per-lane kernels four times the code); AMDGPU runs
# uniform branches natively, so the interpreter units skip structurizing them.
UNIFORM_CFG = ["-mllvm", "-structurizecfg-skip-uniform-regions=true"]

# (source, extra flags)
HIP_UNITS = [
    ("tg_eval_exact.hip", ["-ffp-contract=off", *UNIFORM_CFG]),
    ("tg_eval_fast.hip", ["-ffp-contract=off", *UNIFORM_CFG]),
    ("tg_mesh.hip", ["-ffp-contract=off"]),
    ("tg_capi.hip", ["-ffp-contract=off"]),
    ("tg_export.hip", ["-ffp-contract=off"]),
    ("tg_prune.hip", ["-ffp-contract=off"]),
    ("tg_trace.hip", ["-ffp-contract=off"]),
    ("tg_mlp.hip", ["-ffp-contract=off"]),
    ("tg_surrogate.hip", ["-ffp-contract=off"]),
]

HOST_UNITS = ["tg_tree.cpp", "tg_program.cpp"]
HOST_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-I" + str(CSRC)]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(
        line for line in proc.stdout.splitlines()
        if "packed-fp32-ops" not in line and "not a recognized feature" not in line
    )
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")


def build(jobs: int | None = None, verbose: bool = False) -> Path:
    if not Path(HIPCC).exists():
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    OBJ.mkdir(exist_ok=True)
    jobs = jobs or min(8, int(os.environ.get("MAX_JOBS", os.cpu_count() or 4)))
    tasks = []
    for src, extra in HIP_UNITS:
        obj = OBJ / (src + ".o")
        tasks.append((obj, [HIPCC, *HIP_COMMON, *extra, "-c", str(CSRC / src), "-o", str(obj)]))
    cxx = shutil.which("g++") or "g++"
    for src in HOST_UNITS:
        obj = OBJ / (src + ".o")
        tasks.append((obj, [cxx, *HOST_FLAGS, "-c", str(CSRC / src), "-o", str(obj)]))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = {ex.submit(_run, cmd): obj for obj, cmd in tasks}
        for f in cf.as_completed(futs):
            f.result()
            if verbose:
                print("built", futs[f].name, flush=True)
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *[str(o) for o, _ in tasks]])
    os.replace(tmp, LIB)
    return LIB


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--jobs", type=int, default=None)
    args = ap.parse_args(argv)
    lib = build(args.jobs, verbose=True)
    print(lib)
    return 0


if __name__ == "__main__":
    sys.exit(main())
