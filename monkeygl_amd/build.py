"""Builds the native libraries in-tree.

* ``monkeygl_amd/libmonkeygl_hip.so``: the HIP kernels and the C ABI declared
  in ``include/monkeygl_hip.h`` (hipcc, gfx950 only).
* ``oracle/_build/libmk_oracle.so``: the CPU restatement used by tests and the
  CPU baseline (g++ with OpenMP). Built here because the checker has to exist
  wherever the tests run; the product never loads it.

Run ``python -m monkeygl_amd.build`` or call :func:`build_all`.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shlex
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "monkeygl_amd"
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libmonkeygl_hip.so"
ORACLE_DIR = ROOT / "oracle"
ORACLE_LIB = ORACLE_DIR / "_build" / "libmk_oracle.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-I", str(ROOT / "include"),
]
# files whose arithmetic must match the host restatement bit for bit
NO_CONTRACT = {"mk_plane.hip", "mk_mc.hip", "mk_batch_plane.hip", "mk_overlay.hip", "mk_text.hip"}
# files whose kernels must agree bit for bit with each other (one- and two-pass
# VR, plain and batched): multiply-adds are fused only within one source
# expression, never across expressions by the backend
CONTRACT_ON = {"mk_render.hip", "mk_batch.hip"}

HIP_SOURCES = ["mk_volume.hip", "mk_render.hip", "mk_plane.hip", "mk_mc.hip", "mk_fit.hip", "mk_batch.hip",
               "mk_batch_plane.hip", "mk_encode.hip", "mk_overlay.hip", "mk_text.hip"]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed ({proc.returncode}): {shlex.join(cmd)}\n{proc.stdout}")
    # surface real warnings, drop the host pass's note about the GPU-only feature flag
    lines = [ln for ln in proc.stdout.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
    if lines:
        sys.stderr.write("\n".join(lines) + "\n")


def _hip_object(src: str) -> Path:
    out = OBJ / (Path(src).stem + ".o")
    flags = list(HIP_FLAGS)
    if src in NO_CONTRACT:
        flags.append("-ffp-contract=off")
    elif src in CONTRACT_ON:
        flags.append("-ffp-contract=on")
    _run([HIPCC, *flags, "-c", str(CSRC / src), "-o", str(out)])
    return out


def build_hip(jobs: int = 4) -> Path:
    OBJ.mkdir(parents=True, exist_ok=True)
    srcs = [s for s in HIP_SOURCES if (CSRC / s).exists()]
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_hip_object, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)])
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> Path:
    ORACLE_LIB.parent.mkdir(parents=True, exist_ok=True)
    tmp = ORACLE_LIB.with_suffix(".so.tmp")
    _run(["g++", "-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared",
          "-std=c++17", str(ORACLE_DIR / "mk_oracle.cpp"), "-o", str(tmp)])
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


def build_ref() -> Path | None:
    """Builds ``oracle/_ref/libmkref.so`` (the reference's own sources plus the
    harness, by ``oracle/build_ref.sh``) where the reference is mounted; the
    built library then travels with the tree. Returns None elsewhere."""
    ref = Path(os.environ.get("MONKEYGL_REFERENCE", "/root/reference")) / "core"
    if not (ref / "TransferFunction.cpp").exists():
        return None
    _run(["sh", str(ORACLE_DIR / "build_ref.sh")])
    return ORACLE_DIR / "_ref" / "libmkref.so"


def build_all(jobs: int = 4) -> None:
    with cf.ThreadPoolExecutor(max_workers=3) as ex:
        f1 = ex.submit(build_hip, jobs)
        f2 = ex.submit(build_oracle)
        f3 = ex.submit(build_ref)
        f1.result()
        f2.result()
        f3.result()


if __name__ == "__main__":
    build_all()
    print(f"built {LIB} and {ORACLE_LIB}")
