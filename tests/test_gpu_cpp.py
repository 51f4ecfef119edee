"""The C++ facade (include/ligmg_amd.hpp) used the way the reference's
example uses ligmg::Solver (examples/main.cpp), on its own fixture."""
import os
import subprocess

import pytest

from ligmg_amd import _ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

pytestmark = pytest.mark.gpu


def test_cpp_example_on_reference_fixture(tmp_path):
    exe = tmp_path / "example"
    libdir = os.path.dirname(_ffi.LIB_SOLVER)
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "example.cpp"), "-o", str(exe),
                    f"-L{libdir}", "-lligmg_amd", f"-Wl,-rpath,{libdir}"], check=True)
    p = subprocess.run([str(exe), os.path.join(ROOT, "tests", "golden", "my_matrix.mtx"), str(tmp_path / "dump")],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    print(p.stdout)
    assert p.returncode == 0, p.stdout
    assert "full residual" in p.stdout
    assert "reloaded hierarchy: same solution" in p.stdout
    assert "StrengthMetric,Redistribute" in p.stdout and ",my_matrix,ligmg," in p.stdout
