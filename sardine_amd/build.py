"""Build recipe of the native library sardine_amd/libsardine_hip.so.

hipcc compiles every HIP source for gfx950 (the kernels and the host side of
the C ABI), one object per source in parallel, and links them into a shared
library next to this file.
Flags: -ffp-contract=off keeps the reference's separately rounded
multiplies and adds; packed-FP32 instructions are disabled.

Host sanitizers: `python -m sardine_amd.build --sanitize` builds
sardine_amd/libsardine_hip_san.so with AddressSanitizer and
UndefinedBehaviorSanitizer on the host side of every source (the C ABI,
the TIFF/GeoTIFF/LZW/Deflate/PackBits readers, the geoid parsers, the
server thread); device code is built as usual.  Load it with
SARDINE_LIB=<path> and the sanitizer runtime preloaded:

    SARDINE_LIB=sardine_amd/libsardine_hip_san.so \
    LD_PRELOAD=$(python -m sardine_amd.build --print-asan-runtime) \
    ASAN_OPTIONS=detect_leaks=0,protect_shadow_gap=0 UBSAN_OPTIONS=print_stacktrace=1,halt_on_error=1 \
    python -m pytest tests -m "not gpu"

(detect_leaks=0: the Python interpreter itself is not instrumented and
keeps its allocations to the end; protect_shadow_gap=0: the ROCm runtime
maps memory the default shadow layout reserves.)
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
SAN_LIB = HERE / "libsardine_hip_san.so"
# host-only sanitizer flags: -Xarch_host keeps them off the gfx950 device code
SAN_FLAGS = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
             "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-fno-sanitize-recover=undefined"]
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


def asan_runtime() -> str:
    """Path of the AddressSanitizer runtime the host compiler links against."""
    clang = pathlib.Path(hipcc()).resolve().parent.parent / "lib" / "llvm" / "bin" / "clang"
    if not clang.exists():
        clang = pathlib.Path(shutil.which("clang") or "clang")
    r = subprocess.run([str(clang), "-print-file-name=libclang_rt.asan-x86_64.so"], capture_output=True, text=True)
    return r.stdout.strip()


def build(verbose: bool = False, sanitize: bool = False) -> pathlib.Path:
    """Compile each source to an object in parallel, then link the library
    (with sanitize, the host-instrumented library SAN_LIB instead)."""
    from concurrent.futures import ThreadPoolExecutor

    extra = SAN_FLAGS if sanitize else []
    out = SAN_LIB if sanitize else LIB
    with tempfile.TemporaryDirectory(prefix="sardine_build_") as td:
        objs = [pathlib.Path(td) / (pathlib.Path(s).stem + ".o") for s in SOURCES]
        cmds = [[hipcc(), *flags(), *extra, "-c", str(CSRC / s), "-o", str(o)] for s, o in zip(SOURCES, objs)]
        workers = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
        with ThreadPoolExecutor(workers) as ex:
            list(ex.map(lambda c: _run(c, verbose), cmds))
        tmp = out.with_suffix(".so.tmp")
        # the sanitizer runtime is not linked in: it is preloaded into the
        # (uninstrumented) interpreter, which resolves the library's hooks
        _run([hipcc(), "--offload-arch=gfx950", "-shared", "-o", str(tmp), *map(str, objs), "-lz"], verbose)
        os.replace(tmp, out)
    return out


if __name__ == "__main__":
    if "--print-asan-runtime" in sys.argv:
        print(asan_runtime())
    else:
        print(build(verbose=True, sanitize="--sanitize" in sys.argv))
