"""The C-ABI library loads and exports every entry point include/ripley_hip.h
declares (CPU only: nothing here touches a device)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ripley_hip.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rg_[a-zA-Z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "rg_solve" in names and "rg_matrix_assemble" in names and len(names) >= 20


def test_library_exports_every_declared_symbol():
    from ripley_hip import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libripley_hip.so not built: run __graft_entry__.build()")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(_lib.EXPORTS) == declared()


def test_host_only_entry_points():
    from ripley_hip import _lib
    L = _lib.load()
    assert b"gfx950" in L.rg_version()
    o = _lib.RgOptions()
    L.rg_options_default(ctypes.byref(o))
    assert o.tolerance == 1e-8 and o.iter_max == 100000 and o.sweeps == 1 and o.method == 0


def test_cpp_header_compiles_against_library(tmp_path):
    import shutil
    import subprocess
    from ripley_hip import _lib
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    src = tmp_path / "use.cpp"
    src.write_text('#include "ripley_hip.hpp"\n'
                   'int main() { ripley_hip::SolverOptions o; return o.iter_max == 100000 ? 0 : 1; }\n')
    libdir = os.path.dirname(_lib.LIB_PATH)
    exe = tmp_path / "use"
    subprocess.check_call([gxx, "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-L", libdir,
                           "-lripley_hip", "-Wl,-rpath," + libdir, "-o", str(exe)])
    assert subprocess.call([str(exe)]) == 0


def test_header_constants_match_the_python_mirror():
    # the SpMM shape table of the many-RHS solver: one entry per shape the
    # header counts, the z-marching ones needing all 64 columns
    from ripley_hip import batch as RB
    text = open(HEADER).read()
    shapes = int(re.search(r"#define\s+RG_MULTI_SHAPES\s+(\d+)", text).group(1))
    assert len(RB.SPMM_SHAPES) == shapes
    assert all(RB.SPMM_SHAPES[j][1] == 64 for j in RB.ZMARCH_SHAPES)
    assert all(RB.SPMM_SHAPES[j][1] < 64 for j in range(shapes) if j not in RB.ZMARCH_SHAPES)
