"""Compile the HIP sources of the library into visionaray_amd/libvisionaray_amd.so
for gfx950 with hipcc.  Run as a script or call build()."""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libvisionaray_amd.so"

ARCH = os.environ.get("VR_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: every float operation rounds as written (bit parity with the
# reference arithmetic); no packed-FP32 instructions (-packed-fp32-ops).
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def sources() -> list[Path]:
    return sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed with status {res.returncode}: {cmd[-1]}")


def build(verbose: bool = False) -> Path:
    """Compile every source to an object (in parallel: each translation unit
    launches only its own kernels, so no relocatable device code is needed),
    then link the shared library."""
    from concurrent.futures import ThreadPoolExecutor
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    srcs = sources()
    objs = [objdir / (src.name + ".o") for src in srcs]
    compile_flags = [f for f in HIP_FLAGS if f != "-shared"]
    jobs = max(1, min(len(srcs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    with ThreadPoolExecutor(jobs) as ex:
        futs = [ex.submit(_run, [hipcc, *compile_flags, "-c", "-o", str(o), str(src)], verbose)
                for src, o in zip(srcs, objs)]
        for f in futs:
            f.result()
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, objs)], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
