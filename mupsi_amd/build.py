"""Build recipe for the HIP library (gfx950) and, as test infrastructure, the CPU oracle.

`build_library()` compiles mupsi_amd/csrc/*.{hip,cpp} with hipcc into
mupsi_amd/libmupsi_hip.so (in-tree, so the built file travels with the
repository snapshot).  `build_oracle()` runs oracle/Makefile.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "mupsi_amd"
CSRC = PKG / "csrc"
LIB = PKG / "libmupsi_hip.so"
ARCH = "gfx950"

SOURCES = ["kernels.hip", "guide.hip", "scene_host.cpp", "host_bvh.cpp", "scene_json.cpp"]

# -ffp-contract=off: the only fused multiply-adds are explicit fmaf() calls, which
# fix the float operation order the bit-exact paths rely on.
# -packed-fp32-ops disabled: packed FP32 forms are not used on this hardware pool.
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++20",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP library needs ROCm's hipcc to build")


def build_library(verbose: bool = False) -> Path:
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIP_FLAGS, "-o", str(tmp), "-x", "hip", *[str(CSRC / s) for s in SOURCES]]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if verbose or res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}): {' '.join(cmd)}")
    os.replace(tmp, LIB)
    return LIB


def build_cli(verbose: bool = False) -> Path:
    """`mupsi [scene.json] [out.png]` linked against the in-tree library."""
    exe = PKG / "mupsi"
    tmp = PKG / "mupsi.tmp"
    cmd = ["g++", "-O2", "-std=c++20", "-o", str(tmp), str(CSRC / "cli_main.cpp"),
           f"-L{PKG}", "-lmupsi_hip", "-lz", "-Wl,-rpath,$ORIGIN"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if verbose or res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
    if res.returncode != 0:
        raise RuntimeError(f"CLI build failed: {' '.join(cmd)}")
    os.replace(tmp, exe)
    return exe


def build_oracle(with_reference: bool = True) -> None:
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)
    if with_reference and Path("/root/reference/src/math/random.h").exists():
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "ref"], check=True)


if __name__ == "__main__":
    build_oracle()
    print(build_library(verbose="-v" in sys.argv))
    print(build_cli())
