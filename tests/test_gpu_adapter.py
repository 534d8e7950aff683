"""The C++ host adapter (nereus_cfd_amd/host/PressureBackendHIP.h) driven by a C++ program that restates the
reference's pressure_dirichlet_check (tests/suite_pressure.cpp:79-232): compiled with g++ against libnppe.so and
run on the GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.gpu
def test_cpp_adapter_dirichlet_check(tmp_path):
    exe = str(tmp_path / "test_adapter")
    lib = os.path.join(ROOT, "nereus_cfd_amd")
    subprocess.run(["g++", "-std=c++17", "-O2", os.path.join(ROOT, "tests", "cpp", "test_adapter.cpp"), "-o", exe,
                    "-L" + lib, "-lnppe", "-Wl,-rpath," + lib], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASSED" in r.stdout


def test_cpp_adapter_compiles():
    """CPU: the adapter header and the test program compile and link against the library."""
    out = os.path.join(ROOT, "nereus_cfd_amd", "_build", "test_adapter_cpu")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    lib = os.path.join(ROOT, "nereus_cfd_amd")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", os.path.join(ROOT, "tests", "cpp",
                    "test_adapter.cpp"), "-o", out, "-L" + lib, "-lnppe", "-Wl,-rpath," + lib], check=True)
