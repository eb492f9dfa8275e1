"""Host build of the device elementary functions (stepfactor.h): the
reciprocal division div_rn must equal IEEE division and the step-size
factors must be the correctly rounded powers (no GPU needed)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_device_elementary_functions_on_host(tmp_path):
    exe = tmp_path / "stepfactor_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-fopenmp", "-ffp-contract=off",
                    "-I", os.path.join(ROOT, "nest-gpu_amd", "csrc"),
                    os.path.join(ROOT, "tests", "host", "stepfactor_check.cpp"),
                    "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "0 mismatches" in r.stdout
