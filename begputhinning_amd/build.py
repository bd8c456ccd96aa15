"""Build the HIP library (libbgt.so, gfx950) in-tree with hipcc.

    python -m begputhinning_amd.build          # library only
    python -m begputhinning_amd.build --all    # library + oracle (+ reference build)

The oracle and the reference build are test infrastructure (oracle/Makefile);
the library never links them.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libbgt.so")
ARCH = os.environ.get("BGT_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["thin2d.hip", "thin3d.hip", "sparse3d.hip", "tile3d.hip", "batch.hip", "stream2d.hip", "seggraph.hip", "capi.cpp", "segments.cpp", "synth.cpp",
           "server.cpp"]

COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
          "-I" + os.path.join(ROOT, "include")]
# packed-FP32 instructions are disabled for all device code (no float math on
# this path, and the compiler must not form them)
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def _compile(src, obj_dir=OBJ, extra=()):
    path = os.path.join(CSRC, src)
    obj = os.path.join(obj_dir, src + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC, "-c", path, "-o", obj] + COMMON + DEVICE + list(extra)
    else:
        # host-only translation units: hipcc as the host compiler, no device code
        cmd = [HIPCC, "-c", "-x", "c++", path, "-o", obj, "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
               "-ffp-contract=off"] + COMMON + list(extra)
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in l)
    return src, r.returncode, out, cmd


def build_library(verbose=False, out_dir=None, host_flags=()):
    """Builds libbgt.so in-tree, or into out_dir with host_flags added to
    every compile and to the link (each flag is passed to the host
    compilation only: a sanitizer build of the host code, say)."""
    obj_dir = os.path.join(out_dir, "_build") if out_dir else OBJ
    lib_path = os.path.join(out_dir, "libbgt.so") if out_dir else LIB
    extra = [a for f in host_flags for a in ("-Xarch_host", f)]
    os.makedirs(obj_dir, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")))
    failed = []
    with cf.ThreadPoolExecutor(jobs) as ex:
        for src, rc, out, cmd in ex.map(lambda x: _compile(x, obj_dir, extra), srcs):
            if out.strip() and (verbose or rc):
                print(f"[{src}]\n{out}", file=sys.stderr)
            if rc:
                failed.append(src)
    if failed:
        raise RuntimeError("hipcc failed for: " + ", ".join(failed))
    objs = [os.path.join(obj_dir, s + ".o") for s in srcs]
    tmp = lib_path + ".tmp"
    cmd = [HIPCC, "-shared", "-o", tmp] + objs + ["--offload-arch=" + ARCH, "-fPIC"] + extra
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, lib_path)
    return lib_path


def build_oracle():
    make = shutil.which("make")
    if not make:
        raise RuntimeError("make not found")
    subprocess.run([make, "-s", "-C", os.path.join(ROOT, "oracle"), "oracle"], check=True)
    if os.path.isdir("/root/reference"):
        subprocess.run([make, "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)


def build_dropin_drivers():
    """Test drivers written against the drop-in C++ headers (include/2D,
    include/3D), linked to libbgt.so: tests/dropin/dropin2d and dropin3d."""
    src = os.path.join(ROOT, "tests", "dropin", "dropin_main.cpp")
    if not os.path.exists(src):
        return []
    cxx = shutil.which("g++") or shutil.which("c++")
    if not cxx:
        raise RuntimeError("no C++ compiler for the drop-in drivers")
    outs = []
    for dim in (2, 3):
        exe = os.path.join(ROOT, "tests", "dropin", f"dropin{dim}d")
        cmd = [cxx, "-std=c++14", "-O1", "-Wall", "-Wextra", f"-DDIM={dim}",
               "-I" + os.path.join(ROOT, "include"), src, "-L" + HERE, "-lbgt",
               "-Wl,-rpath,$ORIGIN/../../begputhinning_amd", "-o", exe]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode:
            raise RuntimeError("drop-in driver build failed:\n" + r.stdout + r.stderr)
        outs.append(exe)
    return outs


def main(argv):
    build_library(verbose="-v" in argv)
    build_dropin_drivers()
    if "--all" in argv:
        build_oracle()
    print(LIB)


if __name__ == "__main__":
    main(sys.argv[1:])
