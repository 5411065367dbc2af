"""Build recipe of the native library sardine_amd/libsardine_hip.so.

hipcc compiles every HIP source for gfx950 (the kernels and the host side of
the C ABI), one object per source in parallel, and links them into a shared
library next to this file.
Flags: -ffp-contract=off keeps the reference's separately rounded
multiplies and adds; packed-FP32 instructions are disabled.
"""
from __future__ import annotations

import os
import pathlib
import shutil
import subprocess
import sys
import tempfile

HERE = pathlib.Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "libsardine_hip.so"
SOURCES = ["abi.hip", "tc.hip", "tc32.hip", "tc_batch.hip", "cal.hip", "speckle.hip", "misc.hip", "export.hip", "multilook.hip", "dem_io.hip"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and pathlib.Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def flags() -> list:
    return ["-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-ffp-contract=off",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-pass-failed", "-Wno-unused-command-line-argument"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = ("packed-fp32-ops", "warning: unknown target feature")
    err = "\n".join(l for l in r.stderr.splitlines() if not any(n in l for n in noise))
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed ({r.returncode}):\n{err}")
    if verbose and err.strip():
        print(err, file=sys.stderr)


def build(verbose: bool = False) -> pathlib.Path:
    """Compile each source to an object in parallel, then link the library."""
    from concurrent.futures import ThreadPoolExecutor

    with tempfile.TemporaryDirectory(prefix="sardine_build_") as td:
        objs = [pathlib.Path(td) / (pathlib.Path(s).stem + ".o") for s in SOURCES]
        cmds = [[hipcc(), *flags(), "-c", str(CSRC / s), "-o", str(o)] for s, o in zip(SOURCES, objs)]
        workers = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
        with ThreadPoolExecutor(workers) as ex:
            list(ex.map(lambda c: _run(c, verbose), cmds))
        tmp = LIB.with_suffix(".so.tmp")
        _run([hipcc(), "--offload-arch=gfx950", "-shared", "-o", str(tmp), *map(str, objs), "-lz"], verbose)
        os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
