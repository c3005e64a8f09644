"""Builds libb64hip.so (gfx950) in-tree with hipcc.

    python -m binius64_amd.build          # compile
The library lands next to this file so it travels with the source tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libb64hip.so"
SOURCES = ["api.hip", "kernels.hip", "batch.hip", "server.hip", "pcs.hip", "vision.hip"]
ARCH = os.environ.get("B64_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def flags() -> list[str]:
    return [
        "-O3", "-std=c++17", "-fPIC", "-shared",
        f"--offload-arch={ARCH}",
        # never emit packed-FP32 instructions in device code
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
        "-I", str(PKG.parent / "include"),
    ]


OBJ_DIR = PKG / "_build"


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed")
    # keep the host-pass warning about the device-only target feature out of the way
    noise = [l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()]
    if verbose and noise:
        sys.stderr.write("\n".join(noise) + "\n")


def build(verbose: bool = False) -> Path:
    """Compiles every source to an object in parallel (no device code crosses a
    source file, so no relocatable device code is needed), then links."""
    from concurrent.futures import ThreadPoolExecutor

    OBJ_DIR.mkdir(exist_ok=True)
    compile_flags = [f for f in flags() if f != "-shared"]
    objs = [OBJ_DIR / (Path(src).stem + ".o") for src in SOURCES]
    jobs = int(os.environ.get("MAX_JOBS", "0")) or (os.cpu_count() or 4)
    with ThreadPoolExecutor(max_workers=max(1, min(len(SOURCES), jobs))) as ex:
        futs = [ex.submit(_run, [hipcc(), *compile_flags, "-c", str(CSRC / src), "-o", str(obj)], verbose)
                for src, obj in zip(SOURCES, objs)]
        for f in futs:
            f.result()
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", *[str(o) for o in objs], "-o", str(tmp)], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
