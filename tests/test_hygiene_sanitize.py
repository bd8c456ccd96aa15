"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer, no GPU
needed.

1. The library's tracer, centerline post-processing and MEX column layout
   (csrc/segments.cpp) and the oracle (oracle/oracle.c) compiled with
   -fsanitize=address,undefined into one program that checks them against
   each other on seeded images and volumes (tests/hygiene/sanitize_host.cpp).
2. The whole library built with its host code sanitized (device code as
   always), and the CPU tests of its C-ABI -- argument checks, exported
   symbols, the drop-in C++ headers' drivers and their failure paths without
   a GPU, the host tracer, the MEX layout, the allocation hook -- run against
   it with the sanitizer runtime preloaded.  A sanitizer report fails the
   run (-fno-sanitize-recover)."""
import glob
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address", "-fsanitize=undefined", "-fno-sanitize-recover=all",
       "-fno-omit-frame-pointer"]


def _gcc_has_sanitizers():
    gcc = shutil.which("gcc")
    if not gcc:
        return False
    r = subprocess.run([gcc, "-print-file-name=libasan.so"], capture_output=True, text=True)
    return r.returncode == 0 and os.path.isabs(r.stdout.strip())


@pytest.mark.skipif(not _gcc_has_sanitizers(), reason="no gcc with sanitizer runtimes")
def test_tracer_and_oracle_under_sanitizers(tmp_path):
    inc = ["-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "include")]
    o1 = str(tmp_path / "oracle.o")
    o2 = str(tmp_path / "segments.o")
    exe = str(tmp_path / "sanitize_host")
    steps = [
        ["gcc", "-std=c11", "-O1", "-g"] + SAN + ["-c", os.path.join(ROOT, "oracle", "oracle.c"), "-o", o1],
        ["g++", "-std=c++17", "-O1", "-g"] + SAN + inc +
        ["-c", os.path.join(ROOT, "begputhinning_amd", "csrc", "segments.cpp"), "-o", o2],
        ["g++", "-std=c++17", "-O1", "-g"] + SAN + inc +
        [os.path.join(ROOT, "tests", "hygiene", "sanitize_host.cpp"), o2, o1, "-pthread", "-o", exe],
    ]
    for cmd in steps:
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-3000:]
    r = subprocess.run([exe, "60"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.startswith("ok"), (r.stdout[-2000:], r.stderr[-4000:])


def _clang_asan_runtime():
    hits = sorted(glob.glob("/opt/rocm/lib/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    return hits[-1] if hits else None


@pytest.mark.skipif(not os.path.exists("/opt/rocm/bin/hipcc") or not _clang_asan_runtime(),
                    reason="no hipcc with the sanitizer runtime")
def test_library_host_paths_under_sanitizers(tmp_path):
    sys.path.insert(0, ROOT)
    from begputhinning_amd import build
    lib_dir = str(tmp_path / "asan")
    os.makedirs(lib_dir)
    build.build_library(out_dir=lib_dir, host_flags=SAN + ["-g"])
    env = dict(os.environ)
    env.update({
        "LD_PRELOAD": _clang_asan_runtime(),
        "ASAN_OPTIONS": "detect_leaks=0:abort_on_error=1",  # the interpreter's own allocations are not checked
        "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1",
        "BGT_LIBRARY": os.path.join(lib_dir, "libbgt.so"),
        "LD_LIBRARY_PATH": lib_dir + os.pathsep + env.get("LD_LIBRARY_PATH", ""),
    })
    tests = ["test_device_args.py", "test_capi_exports.py", "test_dropin_host.py",
             "test_segments_host.py", "test_mex_columns.py", "test_alloc_hook.py"]
    cmd = [sys.executable, "-m", "pytest", "-x", "-q", "-m", "not gpu", "-p", "no:cacheprovider"] + \
          [os.path.join(ROOT, "tests", t) for t in tests if os.path.exists(os.path.join(ROOT, "tests", t))]
    r = subprocess.run(cmd, capture_output=True, text=True, env=env, cwd=str(tmp_path), timeout=900)
    assert r.returncode == 0, (r.stdout[-4000:], r.stderr[-4000:])
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
    import re
    m = re.search(r"(\d+) passed", r.stdout)
    assert m and int(m.group(1)) >= 20, r.stdout[-2000:]  # the tests ran, not skipped
