"""The Julia binding sketched in the integration notes against the C ABI: every
``ccall`` names a function include/qgybj.h declares, with the same number of
arguments and matching C types, and the ``QGConfig`` mirror has the fields of
``struct qg_config`` in the same order with the same types.  (No Julia
toolchain is in the image, so the sketch is checked by reading it.)"""

import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DOC = os.path.join(ROOT, "INTEGRATION.md")

# the document is not shipped with every copy of the tree the tests run from
pytestmark = pytest.mark.skipif(not os.path.exists(DOC), reason="the integration notes are not present")


def _julia_block():
    if not os.path.exists(DOC):
        return ""
    with open(DOC) as fh:
        text = fh.read()
    blocks = re.findall(r"```julia\n(.*?)```", text, re.S)
    mod = [b for b in blocks if "module QGYBJplusHIP" in b]
    assert len(mod) == 1
    return mod[0]


def _header():
    with open(os.path.join(ROOT, "include", "qgybj.h")) as fh:
        text = fh.read()
    return re.sub(r"/\*.*?\*/", "", text, flags=re.S)


def _c_prototypes():
    protos = {}
    for ret, name, args in re.findall(r"^(int|const char\s*\*)\s*(qg_\w+)\s*\(([^)]*)\);", _header(), re.M):
        types = []
        for a in args.split(","):
            a = " ".join(a.split())
            if not a or a == "void":
                continue
            # drop the parameter name: the last identifier
            t = re.sub(r"\s*\b\w+$", "", a).replace(" *", "*")
            types.append(t)
        protos[name] = (ret.replace(" ", ""), types)
    return protos


def _c_config_fields():
    body = re.search(r"typedef struct qg_config \{(.*?)\} qg_config;", _header(), re.S).group(1)
    fields = []
    for decl in body.split(";"):
        decl = " ".join(decl.split())
        if not decl:
            continue
        m = re.match(r"((?:const )?\w+\*?)\s+(.*)", decl.replace(" *", "*"))
        ctype, names = m.group(1), m.group(2)
        for n in names.split(","):
            fields.append((n.strip(), ctype))
    return fields


# Julia type in a ccall -> C types it may stand for
JULIA_TO_C = {
    "Cint": {"int"},
    "Cdouble": {"double"},
    "Clonglong": {"long long"},
    "Csize_t": {"size_t"},
    "Ptr{Cvoid}": {"qg_handle*", "const qg_handle*", "void*", "const void*", "double*", "const double*"},
    "Ptr{ComplexF64}": {"double*", "const double*"},  # interleaved (re, im)
    "Ptr{Cdouble}": {"double*", "const double*"},
    "Ptr{Cint}": {"int*", "const int*"},
    "Ref{Cint}": {"int*"},
    "Ref{Cdouble}": {"double*"},
    "Ref{Csize_t}": {"size_t*"},
    "Ref{QGConfig}": {"const qg_config*"},
    "Ref{Ptr{Cvoid}}": {"qg_handle**"},
}
JULIA_RET = {"Cint": "int", "Cstring": "constchar*"}
CONFIG_TYPES = {"Cint": "int", "Cdouble": "double", "Ptr{Cdouble}": "const double*", "Ptr{Cint}": "const int*"}


def _ccalls(block):
    return re.findall(r"ccall\(\(:(\w+), lib\), (\w+), \(([^()]*)\)", block)


def test_sketch_has_ccalls():
    calls = _ccalls(_julia_block())
    names = {c[0] for c in calls}
    for needed in ("qg_create", "qg_destroy", "qg_step", "qg_run", "qg_check", "qg_set_field", "qg_get_field",
                   "qg_last_error", "qg_invert_q_to_psi", "qg_invert_B_to_A", "qg_compute_velocities"):
        assert needed in names, needed


@pytest.mark.parametrize("call", _ccalls(_julia_block()), ids=lambda c: c[0])
def test_ccall_matches_header(call):
    name, ret, args = call
    protos = _c_prototypes()
    assert name in protos, f"{name} is not declared in include/qgybj.h"
    cret, ctypes = protos[name]
    assert JULIA_RET[ret] == cret, (name, ret, cret)
    jtypes = [a.strip() for a in args.split(",") if a.strip()]
    assert len(jtypes) == len(ctypes), (name, jtypes, ctypes)
    for j, c in zip(jtypes, ctypes):
        assert j in JULIA_TO_C, f"{name}: no C mapping for Julia type {j}"
        assert c in JULIA_TO_C[j], (name, j, c)


def test_qgconfig_mirror_matches_struct():
    block = _julia_block()
    body = re.search(r"struct QGConfig\n(.*?)\nend", block, re.S).group(1)
    body = re.sub(r"#.*", "", body)
    jfields = [(n, t) for n, t in re.findall(r"(\w+)::([\w{}]+)", body)]
    cfields = _c_config_fields()
    assert [n for n, _ in jfields] == [n for n, _ in cfields]
    for (n, jt), (_, ct) in zip(jfields, cfields):
        assert CONFIG_TYPES[jt] == ct, (n, jt, ct)


def test_sketch_positional_config_has_every_field():
    """hip_model's QGConfig(...) call passes one value per field."""
    block = re.sub(r"#[^\n]*", "", _julia_block())  # comments may hold commas
    start = block.index("cfg = QGConfig(")
    depth, i = 0, start + len("cfg = QGConfig")
    args, cur = [], ""
    while True:
        ch = block[i]
        if ch in "([{":
            depth += 1
            if depth == 1:
                i += 1
                continue
        elif ch in ")]}":
            depth -= 1
            if depth == 0:
                args.append(cur)
                break
        if ch == "," and depth == 1:
            args.append(cur)
            cur = ""
        else:
            cur += ch
        i += 1
    args = [a.strip() for a in args if a.strip()]
    assert len(args) == len(_c_config_fields()), args
