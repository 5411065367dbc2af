"""The host side of the C ABI under AddressSanitizer + UndefinedBehaviorSanitizer.

Builds sardine_amd/libsardine_hip_san.so (sardine_amd/build.py --sanitize:
host code instrumented with -Xarch_host flags, device code as usual) and
reruns the CPU tests that drive the host code -- the C-ABI argument checks,
layouts, the TIFF/GeoTIFF/BigTIFF/COG writers and readers with crafted
headers, LZW/Deflate/PackBits decoding, the WW15MGH and binary geoid
parsers, error payloads -- in a child interpreter with the sanitizer
runtime preloaded.  Any sanitizer report fails the child run.
"""
import os
import pathlib
import subprocess
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def san_lib():
    from sardine_amd import build

    try:
        build.hipcc()
    except RuntimeError:
        pytest.skip("hipcc not available")
    return build.build(sanitize=True), build.asan_runtime()


def test_cpu_suite_under_asan_ubsan(san_lib):
    lib, runtime = san_lib
    assert pathlib.Path(runtime).exists(), runtime
    env = dict(os.environ, SARDINE_LIB=str(lib), LD_PRELOAD=runtime,
               ASAN_OPTIONS="detect_leaks=0,protect_shadow_gap=0,abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1,halt_on_error=1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu",
                        "tests/test_abi_cpu.py", "tests/test_oracle_reference.py"],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=1200)
    tail = (r.stdout + r.stderr)[-4000:]
    assert r.returncode == 0, tail
    assert "AddressSanitizer" not in r.stderr and "runtime error:" not in r.stderr, tail
    assert " passed" in r.stdout, tail
