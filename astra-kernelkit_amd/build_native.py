"""Build the gfx950 shared library ``kernelkit/_lib/libkernelkit_hip.so``.

Compiles every ``csrc/*.hip`` with hipcc (ahead of time, no runtime
compilation) and links them into one shared library exporting the C ABI of
``include/kernelkit_hip.h``.  Usage: ``python build_native.py [-j N]``.
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
CSRC = HERE / "csrc"
INCLUDE = ROOT / "include"
OUT_DIR = HERE / "kernelkit" / "_lib"
LIB_NAME = "libkernelkit_hip.so"
ARCH = os.environ.get("KK_OFFLOAD_ARCH", "gfx950")

COMMON_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
    f"-I{INCLUDE}", f"-I{CSRC}",
]


def hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(exe).exists():
        raise RuntimeError("hipcc not found: cannot build the HIP library")
    return exe


def build(jobs: int = 8, verbose: bool = False) -> Path:
    """Compile and link; returns the path of the shared library."""
    sources = sorted(CSRC.glob("*.hip"))
    if not sources:
        raise RuntimeError(f"no HIP sources in {CSRC}")
    obj_dir = HERE / "_build"
    obj_dir.mkdir(exist_ok=True)
    OUT_DIR.mkdir(parents=True, exist_ok=True)
    cc = hipcc()

    def compile_one(src: Path) -> Path:
        obj = obj_dir / (src.stem + ".o")
        cmd = [cc, *COMMON_FLAGS, "-c", str(src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src.name}:\n{res.stderr}")
        if verbose and res.stderr.strip():
            print(res.stderr, file=sys.stderr)
        return obj

    with ThreadPoolExecutor(max_workers=max(1, jobs)) as pool:
        objs = list(pool.map(compile_one, sources))
    tmp = OUT_DIR / (LIB_NAME + ".tmp")
    cmd = [cc, "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs),
           "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    final = OUT_DIR / LIB_NAME
    os.replace(tmp, final)
    return final


def main() -> None:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("-j", "--jobs", type=int, default=8)
    ap.add_argument("-v", "--verbose", action="store_true")
    args = ap.parse_args()
    print(build(args.jobs, args.verbose))


if __name__ == "__main__":
    main()
