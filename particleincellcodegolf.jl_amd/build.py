"""Builds libpicgolf.so (HIP kernels for gfx950 + the C ABI) in-tree.

    python particleincellcodegolf.jl_amd/build.py

The library lands in particleincellcodegolf.jl_amd/lib/.  All device code is
compiled with -ffp-contract=off (the FP64 path must round exactly as the CPU
restatement does) and with packed FP32 instructions disabled.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIB_DIR, "libpicgolf.so")
ARCH = os.environ.get("PIC_OFFLOAD_ARCH", "gfx950")

SOURCES = ["kernels_1d.hip", "kernels_mixed.hip", "kernels_mixed_aux.hip", "kernels_small.hip", "noalloc.hip", "kernels_2d.hip", "kernels_1d2v.hip", "kernels_2d3v.hip", "kernels_2d3v_mixed.hip",
           "kernels_2d3v_lg.hip", "kernels_em1d.hip", "capi.cpp", "capi2d.cpp", "capi1d2v.cpp", "capi2d3v.cpp",
           "capi_em1d.cpp"]


# Per-unit compiler options.  The mixed midpoint push schedules better with
# LLVM's iterative ILP machine scheduler: it averaged 93.5 against 95.2 us
# over five alternating runs (instruction order only; the results are the
# same bit for bit).  The mixed Simpson push ran about 5% slower under it,
# so it lives in a unit of its own (kernels_mixed_aux.hip) without it.
EXTRA_FLAGS = {"kernels_mixed.hip": ["-mllvm", "-amdgpu-sched-strategy=iterative-ilp"]}


def hipcc():
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


OBJ_DIR = os.path.join(LIB_DIR, "obj")
SAN_LIB = os.path.join(LIB_DIR, "san", "libpicgolf.so")
# AddressSanitizer + UndefinedBehaviorSanitizer on the host code only (each
# -fsanitize= behind -Xarch_host: the device code is built as always)
HOST_SANITIZE = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
                 "-Xarch_host", "-fno-sanitize-recover=all", "-Xarch_host", "-fno-omit-frame-pointer"]


def build(verbose=False):
    os.makedirs(OBJ_DIR, exist_ok=True)
    objs = []
    common = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall",
              "-Wno-unused-function", "-I", os.path.join(HERE, "..", "include")]
    cmds = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc(), *common, "-c", path, "-o", obj]
        # hipcc compiles .cpp as HIP too: every unit gets the same target flags
        cmd[1:1] = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
        cmd[1:1] = EXTRA_FLAGS.get(src, [])
        cmds.append(cmd)
        objs.append(obj)
    # the units are independent: compile them side by side
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        for f in [ex.submit(run, cmd, verbose) for cmd in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    rocm_lib = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib")
    run([hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs, f"-L{rocm_lib}",
         "-lrocprofiler-sdk-roctx", f"-Wl,-rpath,{rocm_lib}"], verbose)
    os.replace(tmp, LIB)
    return LIB


def build_sanitized(verbose=False):
    """libpicgolf with the host runtime (the C ABI, capi*.cpp) built under
    ASan + UBSan into lib/san/, linked with the kernels of the last build().
    It needs the clang ASan runtime preloaded (tests/test_sanitizers_cpu.py)."""
    if not all(os.path.exists(os.path.join(OBJ_DIR, os.path.splitext(f)[0] + ".o")) for f in SOURCES):
        build(verbose)
    san_dir = os.path.dirname(SAN_LIB)
    os.makedirs(san_dir, exist_ok=True)
    common = ["-O1", "-g", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
              "-I", os.path.join(HERE, "..", "include")]
    cmds, objs = [], []
    for src in SOURCES:
        stem = os.path.splitext(src)[0]
        if not src.endswith(".cpp"):
            objs.append(os.path.join(OBJ_DIR, stem + ".o"))
            continue
        obj = os.path.join(san_dir, stem + ".o")
        cmds.append([hipcc(), f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                     *HOST_SANITIZE, *common, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    with concurrent.futures.ThreadPoolExecutor(max(1, min(len(cmds), os.cpu_count() or 1))) as ex:
        for f in [ex.submit(run, cmd, verbose) for cmd in cmds]:
            f.result()
    rocm_lib = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib")
    run([hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", SAN_LIB, *objs, f"-L{rocm_lib}",
         "-lrocprofiler-sdk-roctx", f"-Wl,-rpath,{rocm_lib}"], verbose)
    return SAN_LIB


def asan_runtime():
    """The clang ASan runtime matching hipcc (to preload for SAN_LIB)."""
    r = subprocess.run([hipcc(), "-print-file-name=libclang_rt.asan-x86_64.so"], capture_output=True, text=True)
    return r.stdout.strip()


def run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    err = "\n".join(l for l in (r.stderr or "").splitlines() if noise not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{err}")
    if verbose and err.strip():
        print(err, file=sys.stderr)


if __name__ == "__main__":
    print(build(verbose=True))
    if len(sys.argv) > 1 and sys.argv[1] == "sanitize":
        print(build_sanitized(verbose=True))
