"""The C++ host boundary end to end on the GPU: tests/cpp/adapter_main.cpp
(built by __graft_entry__.build() into esys-escript.github.io_amd/lib/
rg_cpp_adapter) drives include/ripley_hip.hpp with host arrays only, as an
escript adapter would (RipleyDomain::addToSystemMatrix, insertConstraint,
AbstractSystemMatrix::ypAx / setToSolution), and checks the Q1 patch test and
the nodally exact solution u = z - z^2/2 of -u'' = 1 on a Brick."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "esys-escript.github.io_amd", "lib", "rg_cpp_adapter")


def test_cpp_adapter_program():
    assert os.path.exists(EXE), "rg_cpp_adapter is not built; run __graft_entry__.build()"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("ok  ") == 4
