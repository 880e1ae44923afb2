"""The C ABI boundary without a GPU: the library loads, exports every symbol
include/qgybj.h declares, and the ctypes mirror of qg_config has the C
layout.  No compute call is made."""

import ctypes
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "qgybj.h")
LIB = os.path.join(ROOT, "qgybjplus.jl_amd", "libqgybj.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(qg_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_entry_points():
    names = declared_functions()
    assert "qg_create" in names and "qg_step" in names and "qg_check" in names
    assert len(names) >= 17


def test_binding_lists_every_declared_symbol(qg):
    from importlib import import_module
    lib_mod = import_module(qg.__name__ + "._lib")
    assert sorted(lib_mod.EXPORTED) == declared_functions()


@pytest.mark.skipif(not os.path.exists(LIB), reason="libqgybj.so not built")
def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="libqgybj.so not built")
def test_library_loads_and_resolves_symbols():
    lib = ctypes.CDLL(LIB)
    for n in declared_functions():
        assert getattr(lib, n) is not None


def test_qg_config_layout_matches_c(qg):
    """Compile a probe against the header and compare sizeof/offsetof with the
    ctypes structure the binding passes to qg_create."""
    from importlib import import_module
    lib_mod = import_module(qg.__name__ + "._lib")
    fields = [f for f, _ in lib_mod.QgConfig._fields_]
    body = "\n".join(f'  printf("{f} %zu\\n", offsetof(qg_config, {f}));' for f in fields)
    src = ("#include <stdio.h>\n#include <stddef.h>\n#include \"qgybj.h\"\nint main(void) {\n"
           f'  printf("sizeof %zu\\n", sizeof(qg_config));\n{body}\n  return 0;\n}}\n')
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "probe.c")
        exe = os.path.join(d, "probe")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    got = dict((k, int(v)) for k, v in (line.split() for line in out.splitlines()))
    assert got["sizeof"] == ctypes.sizeof(lib_mod.QgConfig)
    for f in fields:
        assert got[f] == getattr(lib_mod.QgConfig, f).offset, f


def test_header_compiles_as_c_and_cpp():
    with tempfile.TemporaryDirectory() as d:
        for comp, ext in (("gcc", "c"), ("g++", "cpp")):
            p = os.path.join(d, "inc." + ext)
            open(p, "w").write('#include "qgybj.h"\nint main(void) { return QG_OK; }\n')
            subprocess.run([comp, "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), p, "-o",
                            os.path.join(d, "inc_" + ext)], check=True)


@pytest.mark.skipif(not os.path.exists(LIB), reason="libqgybj.so not built")
def test_memory_required_scales_with_members_and_precision(qg):
    from importlib import import_module
    lib_mod = import_module(qg.__name__ + "._lib")
    one = lib_mod.memory_required(nx=64, ny=64, nz=32, precision=lib_mod.PRECISION_FP64)
    four = lib_mod.memory_required(nx=64, ny=64, nz=32, precision=lib_mod.PRECISION_FP64, members=4)
    mixed = lib_mod.memory_required(nx=64, ny=64, nz=32, precision=lib_mod.PRECISION_MIXED)
    npts = 64 * 64 * 32
    assert 300 * npts < one < 600 * npts  # ~18 complex FP64 fields plus work buffers
    assert 3.9 * one < four < 4.1 * one
    assert mixed < one


def test_poisson_trace_matches_its_distribution(qg):
    tr = qg.serving.poisson_trace(4000, rate=2.0, seed=3)
    gaps = np.diff([0.0] + [r.arrival for r in tr])
    assert abs(np.mean(gaps) - 0.5) < 0.03
    assert {r.nx for r in tr} == {64, 128, 256, 384} and {r.nz for r in tr} == {32, 64, 128}
    steps = np.array([r.steps for r in tr])
    assert steps.min() >= 100 and steps.max() <= 2000
    assert abs(np.median(np.log(steps)) - 0.5 * (np.log(100) + np.log(2000))) < 0.1
    assert 0.45 < np.mean([r.stratification == "constant" for r in tr]) < 0.55


def test_bench_scan_shapes_match_the_compiled_table():
    """bench.py names column-solve kernels by (levels per thread, quarter
    columns per tile) from the library's shape table; both lists agree."""
    src = open(os.path.join(ROOT, "qgybjplus.jl_amd", "csrc", "qg_colscan.hpp")).read()
    table = re.search(r"kScanShapes\[\]\[2\]\s*=\s*\{(.*?)\};", src, flags=re.S).group(1)
    compiled = tuple((int(m), int(q)) for m, q in re.findall(r"\{\s*(\d+)\s*,\s*(\d+)\s*\}", table))
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench_for_test", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    assert bench.SCAN_SHAPES == compiled
    assert compiled[bench.SCAN_DEFAULT] == (16, 8)
