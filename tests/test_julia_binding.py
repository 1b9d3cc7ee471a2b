"""The Julia binding (acoustic-disguising_amd/julia/hip.jl), checked
without Julia: every ccall names an entry point that include/adfdtd.h
declares and the library exports, passes as many arguments as the C
prototype takes, and the AdfCpml mirror has adf_cpml's fields in order."""
import os
import re

import pytest

import adfdtd._lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JL = os.path.join(ROOT, "acoustic-disguising_amd", "julia", "hip.jl")
HDR = os.path.join(ROOT, "include", "adfdtd.h")


def split_top(s):
    """Split at commas outside (), {} and []."""
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "({[":
            depth += 1
        elif ch in ")}]":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


def balanced(src, start):
    """The text inside the parentheses opening at src[start]."""
    depth = 0
    for i in range(start, len(src)):
        if src[i] == "(":
            depth += 1
        elif src[i] == ")":
            depth -= 1
            if depth == 0:
                return src[start + 1:i]
    raise ValueError("unbalanced")


def ccalls():
    src = "\n".join(line.split("#", 1)[0] for line in open(JL).read().splitlines())
    out = []
    for m in re.finditer(r"ccall\(", src):
        body = balanced(src, m.end() - 1)
        parts = split_top(body)
        name = re.match(r"\(:(\w+),\s*LIB\)", parts[0]).group(1)
        types = split_top(balanced(parts[2], 0)) if parts[2].startswith("(") else []
        out.append((name, len(types), len(parts) - 3))
    return out


def prototypes():
    src = re.sub(r"/\*.*?\*/", "", open(HDR).read(), flags=re.S)
    protos = {}
    for m in re.finditer(r"(?:int|const char\*)\s+(adf_\w+)\s*\(([^)]*)\)\s*;", src):
        args = m.group(2).strip()
        protos[m.group(1)] = 0 if args in ("", "void") else len(split_top(args))
    return protos


def test_every_ccall_matches_a_declared_exported_entry_point():
    calls = ccalls()
    protos = prototypes()
    lib = L.load()
    assert len(calls) >= 20
    for name, ntypes, nargs in calls:
        assert name in protos, name
        assert hasattr(lib, name), name
        assert ntypes == protos[name], (name, ntypes, protos[name])
        assert nargs == ntypes, (name, nargs, ntypes)


def test_cpml_mirror_field_order():
    src = open(JL).read()
    body = re.search(r"struct AdfCpml\n(.*?)\nend", src, re.S).group(1)
    fields = re.findall(r"(\w+)::", body)
    assert fields == [f[0] for f in L.AdfCpml._fields_]


@pytest.mark.parametrize("name,val", [("ADF_F32", 1), ("ADF_F64", 2), ("ADF_F32_REF", 3)])
def test_dtype_codes(name, val):
    src = open(JL).read()
    m = re.search(r"const ADF_F32, ADF_F64, ADF_F32_REF = Cint\((\d)\), Cint\((\d)\), Cint\((\d)\)", src)
    got = dict(zip(("ADF_F32", "ADF_F64", "ADF_F32_REF"), map(int, m.groups())))
    hdr = int(re.search(rf"#define {name} (\d+)", open(HDR).read()).group(1))
    assert got[name] == hdr == val
