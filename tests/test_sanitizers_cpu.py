"""The oracle and the host runtime under AddressSanitizer and
UndefinedBehaviorSanitizer (CPU only; no report may occur).

- oracle/: `make -C oracle sanitize` builds the restatement with
  -fsanitize=address,undefined; the fast oracle tests run against it in a
  child process with the gcc ASan runtime preloaded.
- the host runtime (the C ABI in csrc/capi*.cpp): built with the host-side
  sanitizers into lib/san/ and driven by tests/test_capi_cpu.py (loading,
  exports, configuration checks of every module) with the clang ASan
  runtime preloaded.
"""
import os
import subprocess
import sys

import pytest

from conftest import ROOT

PKG = os.path.join(ROOT, "particleincellcodegolf.jl_amd")
SAN_ENV = {"ASAN_OPTIONS": "detect_leaks=0:abort_on_error=1:halt_on_error=1",
           "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"}
# the oracle tests that finish in seconds (the long physics runs are left out)
SLOW = ("growth_rate", "frequency", "frequencies", "lattice", "inverts", "momentum_x_thermal", "upper_hybrid")


def child(args, env_extra, timeout):
    env = dict(os.environ)
    env.update(SAN_ENV)
    env.update(env_extra)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", "-m", "not gpu", *args],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=timeout)
    tail = (r.stdout + r.stderr)[-3000:]
    assert "AddressSanitizer" not in tail and "runtime error" not in tail, tail
    assert r.returncode == 0, tail
    return tail


def test_oracle_under_asan_ubsan():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    tests = [os.path.join("tests", f) for f in sorted(os.listdir(os.path.join(ROOT, "tests")))
             if f.startswith("test_oracle_")]
    tail = child([*tests, "-k", " and ".join(f"not {k}" for k in SLOW)],
                 {"LD_PRELOAD": asan, "PICORACLE_LIB": os.path.join(ROOT, "oracle", "_san", "libpicoracle_san.so")},
                 timeout=1200)
    print(tail.strip().splitlines()[-1])


def test_host_runtime_under_asan_ubsan():
    import importlib.util
    spec = importlib.util.spec_from_file_location("picgolf_build", os.path.join(PKG, "build.py"))
    b = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(b)
    lib = b.build_sanitized()
    tail = child([os.path.join("tests", "test_capi_cpu.py")], {"LD_PRELOAD": b.asan_runtime(), "PICGOLF_LIB": lib},
                 timeout=600)
    print(tail.strip().splitlines()[-1])
