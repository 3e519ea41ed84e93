"""The Julia binding (highordermri.jl_amd/julia/HighOrderMRIHIP) against the C ABI.

Julia is not installed in this image, so the binding is checked as text against
include/hoe.h, the declaration it must match:
- the Julia ``ModelDesc`` (and the Python ``_lib.ModelDesc``) has the fields of
  ``hoe_model_desc`` in the same order with the same widths;
- every ``ccall((:hoe_..., libhoe), R, (T...), ...)`` names an exported entry point and
  passes as many arguments as the C prototype has, each of the same kind (pointer,
  32-bit integer, 64-bit integer, double);
- Project.toml names the package and its dependencies.
"""
import re
from pathlib import Path

import pytest

from hoe import _lib

ROOT = Path(__file__).resolve().parent.parent
JL = ROOT / "highordermri.jl_amd" / "julia" / "HighOrderMRIHIP"
HDR = ROOT / "include" / "hoe.h"


def _strip_c_comments(text):
    return re.sub(r"/\*.*?\*/", "", text, flags=re.S)


def _c_kind(ctype):
    t = ctype.strip()
    if "*" in t:
        return "ptr"
    if t.endswith("double"):
        return "f64"
    if "int64_t" in t:
        return "i64"
    if t.endswith("int32_t") or t.endswith("int"):
        return "i32"
    raise AssertionError(f"unmapped C type {ctype!r}")


def _jl_kind(jtype):
    t = jtype.strip()
    if t.startswith(("Ptr{", "Ref{")) or t == "Cstring":
        return "ptr"
    if t == "Float64":
        return "f64"
    if t == "Int64":
        return "i64"
    if t in ("Int32", "Cint"):
        return "i32"
    raise AssertionError(f"unmapped Julia type {jtype!r}")


def _header_struct_fields(name):
    text = _strip_c_comments(HDR.read_text())
    body = re.search(r"typedef struct \{([^{}]*)\}\s*" + name + ";", text).group(1)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        m = re.match(r"((?:const\s+)?[a-z_0-9]+)\s+(.*)", decl)
        base, rest = m.group(1), m.group(2)
        for part in rest.split(","):
            part = part.strip()
            ptr = part.startswith("*")
            fields.append((part.lstrip("*").strip(), base + (" *" if ptr else "")))
    return fields


def _header_prototypes():
    text = _strip_c_comments(HDR.read_text())
    protos = {}
    for m in re.finditer(r"\b[a-z_0-9 \*]+?\b(hoe_[a-z_0-9]+)\s*\(([^)]*)\)\s*;", text):
        args = [a.strip() for a in m.group(2).split(",") if a.strip() and a.strip() != "void"]
        types = []
        for a in args:
            a = re.sub(r"\b[a-z_0-9]+$", "", a.replace("*", " * ")).strip()   # drop the parameter name
            types.append(a)
        protos[m.group(1)] = types
    return protos


def _balanced(text, i):
    """The parenthesised group starting at text[i] == '(' and the index after it."""
    assert text[i] == "("
    depth = 0
    for j in range(i, len(text)):
        if text[j] == "(":
            depth += 1
        elif text[j] == ")":
            depth -= 1
            if depth == 0:
                return text[i + 1:j], j + 1
    raise AssertionError("unbalanced parentheses")


def _split_top(s):
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "({[":
            depth += 1
        elif ch in ")}]":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur)
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur)
    return [x.strip() for x in out]


def _julia_ccalls(text):
    calls = []
    for m in re.finditer(r"ccall\(\(:(hoe_[a-z_0-9]+),\s*libhoe\),", text):
        i = m.end()
        rest = text[i:]
        # return type, then the argument-type tuple
        ret_end = rest.index(",")
        j = i + ret_end + 1
        while text[j].isspace():
            j += 1
        tup, _ = _balanced(text, j)
        calls.append((m.group(1), _split_top(tup)))
    return calls


def _julia_struct(text, name):
    body = re.search(r"struct " + name + r"\n(.*?)\nend", text, flags=re.S).group(1)
    fields = []
    for line in body.splitlines():
        line = line.split("#")[0]
        for decl in line.split(";"):
            decl = decl.strip()
            if decl:
                n, t = decl.split("::")
                fields.append((n.strip(), t.strip()))
    return fields


_ELEM = {"float": {"Float32", "ComplexF32"}, "double": {"Float64", "ComplexF64"},
         "uint8_t": {"UInt8"}, "char": {"UInt8"}}


@pytest.mark.parametrize("jl_name,c_name,py", [("ModelDesc", "hoe_model_desc", "ModelDesc"),
                                              ("ModelDescF64", "hoe_model_desc_f64", "ModelDescF64")])
def test_julia_model_desc_matches_header(jl_name, c_name, py):
    text = (JL / "src" / "HighOrderMRIHIP.jl").read_text()
    jl = _julia_struct(text, jl_name)
    hdr = _header_struct_fields(c_name)
    assert [n for n, _ in jl] == [n for n, _ in hdr]
    for (n, jt), (_, ct) in zip(jl, hdr):
        assert _jl_kind(jt) == _c_kind(ct), (n, jt, ct)
        if _c_kind(ct) == "ptr":
            elem = ct.replace("const", "").replace("*", "").strip()
            assert jt[4:-1] in _ELEM[elem], (n, jt, ct)
    # the Python binding's mirror too
    assert [f[0] for f in getattr(_lib, py)._fields_] == [n for n, _ in hdr]


def test_julia_ccalls_match_prototypes():
    text = (JL / "src" / "HighOrderMRIHIP.jl").read_text()
    protos = _header_prototypes()
    calls = _julia_ccalls(text)
    assert len(calls) >= 12
    seen = set()
    for name, jtypes in calls:
        assert name in protos, name
        assert name in _lib.EXPORTS, name
        ctypes_ = protos[name]
        assert len(jtypes) == len(ctypes_), (name, jtypes, ctypes_)
        for jt, ct in zip(jtypes, ctypes_):
            assert _jl_kind(jt) == _c_kind(ct), (name, jt, ct)
        seen.add(name)
    for required in ("hoe_create", "hoe_forward", "hoe_adjoint", "hoe_normal", "hoe_set_weights",
                     "hoe_set_coefficients", "hoe_grad", "hoe_set_kspha", "hoe_set_kspha_dt", "hoe_destroy"):
        assert required in seen, required


def test_julia_project_file():
    tomli = pytest.importorskip("tomli")
    proj = tomli.loads((JL / "Project.toml").read_text())
    assert proj["name"] == "HighOrderMRIHIP"
    assert re.fullmatch(r"[0-9a-f]{8}-[0-9a-f]{4}-[0-9a-f]{4}-[0-9a-f]{4}-[0-9a-f]{12}", proj["uuid"])
    for dep in ("HighOrderMRI", "LinearOperators", "LinearOperatorCollection", "ChainRulesCore", "AMDGPU"):
        assert dep in proj["deps"], dep
    text = (JL / "src" / "HighOrderMRIHIP.jl").read_text()
    for name in ("explicit_gpu_normal", "normalOperator", "route!", "hip_explicit_grad"):
        assert name in text, name
