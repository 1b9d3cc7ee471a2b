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
