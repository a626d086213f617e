"""fix_round128 (csrc/pic_device.h), the rounding of every exact signed
fixed-point sum (the structured 2D3V charge, the Lorenz-gauge and 1D EM
currents, the mixed kinetic diagnostics), is computed from the double's bits.
This compiles it for the host and checks it against its definition
rint(ldexp(a, q)) on the FPU over every class of double and the grids the
kernels use: the device results then equal those of the FPU formulation."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "particleincellcodegolf.jl_amd", "csrc")


@pytest.mark.skipif(shutil.which("hipcc") is None, reason="hipcc is not installed")
def test_fix_round128_matches_its_definition(tmp_path):
    exe = tmp_path / "fixround_check"
    subprocess.run(["hipcc", "--offload-arch=gfx950", "-O2", "-std=c++17", "-I", CSRC,
                    os.path.join(ROOT, "tests", "cpu_checks", "fixround_check.hip"), "-o", str(exe)],
                   check=True, capture_output=True, timeout=600)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(out.stdout)
    assert out.returncode == 0, out.stdout
    assert " 0 mismatches" in out.stdout
