"""The C-ABI library loads and exports every symbol include/*.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import os
import re

import adfdtd._lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            src = open(os.path.join(inc, fn)).read()
            names |= set(re.findall(r"^\s*(?:int|const char\*)\s+(adf_\w+)\s*\(", src, re.M))
    return names


def test_header_matches_binding_list():
    assert declared_symbols() == set(L.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = L.load()
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_version_without_gpu():
    assert L.load().adf_version() == 1


def test_domain_struct_layout_matches_reference_field_order():
    # src/simulation.jl:25-44: 4 Float64, 3 Int, 9 Float64, Int, Float64
    assert ctypes.sizeof(L.AdfDomain) == 18 * 8
    names = [f[0] for f in L.AdfDomain._fields_]
    assert names == ["tmax", "xmax", "ymax", "zmax", "nx", "ny", "nz", "c0", "r0", "cf", "dx", "dy", "dz",
                     "fact_m0", "fact_m1", "dt", "nt", "z0"]


def test_struct_layouts_match_the_c_header(tmp_path):
    """sizeof/offsetof of every struct of include/adfdtd.h, compiled with gcc,
    equal the ctypes mirrors."""
    import subprocess

    structs = {"adf_domain": L.AdfDomain, "adf_cpml": L.AdfCpml, "adf_request": L.AdfRequest,
               "adf_server_info": L.AdfServerInfo}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "adfdtd.h"', "int main(void) {"]
    for cname, cls in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f in cls._fields_:
            lines.append(f'printf("{cname}.{f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                       check=True).stdout.splitlines())
    for cname, cls in structs.items():
        assert int(got[cname]) == ctypes.sizeof(cls), cname
        for f in cls._fields_:
            assert int(got[f"{cname}.{f[0]}"]) == getattr(cls, f[0]).offset, (cname, f[0])


def test_autotune_cache_roundtrip(tmp_path):
    from adfdtd import autotune

    path = str(tmp_path / "sub" / "tune.json")
    assert autotune._load(path) == {}
    autotune._store(path, "solver:1:8x8x8:p2:m0", {"tile_y": 2, "block_z": 8})
    autotune._store(path, "batch:x", {"persistent": 1})
    data = autotune._load(path)
    assert data["solver:1:8x8x8:p2:m0"]["block_z"] == 8 and data["batch:x"]["persistent"] == 1
