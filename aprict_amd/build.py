"""Ahead-of-time build of libaprict_hip.so (hand-written HIP kernels, gfx950).

``python -m aprict_amd.build`` compiles every ``csrc/*.hip`` to an object with
hipcc and links them into ``aprict_amd/lib/libaprict_hip.so``.  The library is
built in-tree so that it travels with the source tree.
"""

from __future__ import annotations

import concurrent.futures
import os
import pathlib
import shutil
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
LIBNAME = "libaprict_hip.so"
ARCH = os.environ.get("APRICT_OFFLOAD_ARCH", "gfx950")

# -packed-fp32-ops: keep the compiler from forming packed FP32 instructions.
HIPFLAGS = [
    "-O3",
    "-std=c++17",
    f"--offload-arch={ARCH}",
    "-fPIC",
    "-munsafe-fp-atomics",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not pathlib.Path(exe).exists():
        msg = "hipcc not found (ROCm is required to build libaprict_hip.so)"
        raise RuntimeError(msg)
    return exe


def _compile(src: pathlib.Path, obj: pathlib.Path, verbose: bool) -> None:
    cmd = [hipcc(), *HIPFLAGS, "-c", str(src), "-o", str(obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True, check=False)
    if res.returncode != 0:
        msg = f"hipcc failed on {src.name}:\n{res.stdout}\n{res.stderr}"
        raise RuntimeError(msg)
    if verbose and res.stderr.strip():
        print(res.stderr, file=sys.stderr)


def build(verbose: bool = False, jobs: int | None = None) -> pathlib.Path:
    """Compile all kernels and link the shared library; returns its path."""
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    objs = [objdir / (s.stem + ".o") for s in srcs]
    jobs = jobs or min(len(srcs), os.cpu_count() or 1, 8)
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_compile, s, o, verbose) for s, o in zip(srcs, objs)]
        for f in futs:
            f.result()
    out = LIBDIR / LIBNAME
    tmp = LIBDIR / (LIBNAME + ".tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True, check=False)
    if res.returncode != 0:
        msg = f"link failed:\n{res.stdout}\n{res.stderr}"
        raise RuntimeError(msg)
    tmp.replace(out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
