"""GF loaders' scaling chain, mirroring src/greens_io.jl.

`load_gf_ref` (src/greens_io.jl:224-477) and `load_gf_mdd` (:536-617) turn
Green's functions stored in physical units into the GF map the in-loop
extrapolation consumes (adf_solver_set_gf_map): physical -> FDTD-recorded
amplitude (C_p, C_v), resampling onto the synthesis time step, the
recording -> injection bridge e, the Kirchhoff-Helmholtz quadrature
weights 4 pi r^2 / N of both surfaces, the impedance conversion z0 and the
one-way (pin) or split (pos/neg) combinations.

The reference reads its GF files with HDF5.jl; this package has no HDF5
reader (h5py is not a dependency), so the loaders take the file's datasets
and attributes as a mapping (`read_gf_file` reads them from an .npz
container with the same names: datasets p_p, p_v, v_p, v_v, t; attributes
dom_dt, dom_dx, backend, units_convention, empirical_scale).

Layouts: the reference's (nt, n_rec, n_src) GF arrays are numpy
(n_src, n_rec, nt) arrays here, as everywhere in the package; the GF maps
take them in that layout (extrapolation.py).
"""
from __future__ import annotations

import math
from typing import Callable, Mapping, Optional

import numpy as np

from .extrapolation import PFromPin, PVFromPin, PVFromPV
from .reverb import julia_step_range, resample_traces
from .simulation import Domain

EXTRAP_TYPES = ("pv_from_pin", "pv_from_pv", "p_from_pin")


def fib_sphere_area(radius: float, n_points: int) -> float:
    """src/io.jl:131: the area element 4 pi r^2 / N of an N-point sphere."""
    return 4 * math.pi * float(radius) ** 2 / int(n_points)


def scaling_factors(dom: Domain, dom_dx: float, dom_dt: float, dt_in: float, cfg: Optional[dict] = None):
    """`_scaling_factors` (src/greens_io.jl:149-166) -> (C_p, C_v, dx_gf):
    C_p = dx_gf^3 dt_in / (dom_dt^2 c^2), C_v = -rho dx_gf^3 dt_in / dom_dt^2,
    with dx_gf the GF file's dom_dx, or 2 xmax / (n - 1) of cfg["greens"]
    ["grid"] when the file has none (NaN)."""
    if math.isnan(dom_dx):
        if cfg is None or "greens" not in cfg or "grid" not in cfg["greens"]:
            raise ValueError("_scaling_factors: the GF file has no dom_dx and cfg has no greens.grid to derive it")
        g = cfg["greens"]["grid"]
        dx_gf = 2 * float(g["xmax"]) / (int(g["n"]) - 1)
    else:
        dx_gf = float(dom_dx)
    C_p = dx_gf ** 3 * dt_in / (float(dom_dt) ** 2 * dom.c0 ** 2)
    C_v = -dom.r0 * dx_gf ** 3 * dt_in / float(dom_dt) ** 2
    return C_p, C_v, dx_gf


def gf_unit_factors(dom: Domain, dt_save: float):
    """src/greens_io.jl:184-188: (C_p, C_v) for GFs generated on `dom`."""
    C_p = dom.dx ** 3 * dt_save / (dom.dt ** 2 * dom.c0 ** 2)
    C_v = -dom.r0 * dom.dx ** 3 * dt_save / dom.dt ** 2
    return C_p, C_v


def gf_to_physical(state: Mapping, C_p: float, C_v: float) -> dict:
    """src/greens_io.jl:197-207: FDTD-recorded -> physical units (Float32
    kernels divided by Float32(C_p) / Float32(C_v))."""
    cp, cv = np.float32(C_p), np.float32(C_v)
    out = {"p_p": np.asarray(state["p_p"], np.float32) / cp, "p_v": np.asarray(state["p_v"], np.float32) / cv,
           "v_p": np.asarray(state["v_p"], np.float32) / cp, "v_v": np.asarray(state["v_v"], np.float32) / cv}
    return {**out, **{k: state[k] for k in ("t", "src_positions", "rec_positions") if k in state}}


def gf_to_raw_fdtd(loaded: Mapping, C_p: float, C_v: float) -> dict:
    """src/greens_io.jl:214-224: physical -> FDTD-recorded units."""
    cp, cv = np.float32(C_p), np.float32(C_v)
    out = {"p_p": np.asarray(loaded["p_p"], np.float32) * cp, "p_v": np.asarray(loaded["p_v"], np.float32) * cv,
           "v_p": np.asarray(loaded["v_p"], np.float32) * cp, "v_v": np.asarray(loaded["v_v"], np.float32) * cv}
    return {**out, **{k: loaded[k] for k in ("t", "src_positions", "rec_positions") if k in loaded}}


def read_gf_file(path: str) -> dict:
    """The datasets and attributes of a GF file as one dict.  .npz
    containers (numpy, no pickles) are read here; HDF5 files would need
    h5py, which this package does not depend on."""
    if path.endswith(".npz"):
        with np.load(path, allow_pickle=False) as z:
            out = {k: z[k] for k in z.files}
        for k in ("dom_dt", "dom_dx", "empirical_scale"):
            if k in out:
                out[k] = float(out[k])
        for k in ("backend", "units_convention"):
            if k in out:
                out[k] = str(out[k])
        return out
    try:
        import h5py  # noqa: F401
    except ImportError as e:
        raise RuntimeError(f"read_gf_file: {path}: reading HDF5 needs h5py, which is not installed; "
                           "write the datasets to an .npz container instead") from e
    import h5py

    with h5py.File(path, "r") as f:
        out = {k: np.asarray(f[k]) for k in f.keys()}
        out.update({k: (v.decode() if isinstance(v, bytes) else v) for k, v in f.attrs.items()})
    return out


def write_gf_file(path: str, gf: Mapping) -> None:
    """An .npz GF container (datasets and scalar attributes) for
    read_gf_file."""
    np.savez(path, **{k: np.asarray(v) for k, v in gf.items()})


def _ref_datasets(f: Mapping, name: str):
    """`_read_gf_ref_h5` (src/greens_io.jl:489-518): the four kernels as
    Float32, t, dom_dt (falling back to t[2] - t[1]), dom_dx (NaN when
    absent), backend and units_convention."""
    k = {c: np.asarray(f[c], np.float32) for c in ("p_p", "p_v", "v_p", "v_v")}
    t = np.asarray(f["t"], np.float64)
    if "dom_dt" in f:
        dom_dt = float(f["dom_dt"])
    elif t.shape[0] > 1:
        dom_dt = float(t[1] - t[0])
    else:
        raise ValueError(f"load_gf_ref: cannot infer dom_dt ({name})")
    dom_dx = float(f["dom_dx"]) if "dom_dx" in f else float("nan")
    return k, t, dom_dt, dom_dx, str(f.get("backend", "")), str(f.get("units_convention", ""))


def _cubic(a: np.ndarray, t_in0: float, dt_in: float, t_out0: float, dt_out: float, n_out: int,
           scale: float) -> np.ndarray:
    """resample_time(A, tin, tout; kind = :cubic) (src/resample.jl:35-77) on
    the device along the last axis: tin = t_in0 + k dt_in, tout = t_out0 +
    k dt_out, output times scale."""
    return resample_traces(a, t_in0, dt_in, t_out0, dt_out, n_out, amp_scale=scale)


def _linear(a: np.ndarray, t_in0: float, dt_in: float, t_out0: float, dt_out: float, n_out: int,
            scale: float) -> np.ndarray:
    """resample_time(A, tin, tout; kind = :linear): BSpline(Linear()), Flat
    outside, times scale, along the last axis (host)."""
    n = a.shape[-1]
    x = 1.0 + ((t_out0 + dt_out * np.arange(n_out)) - t_in0) / dt_in
    x = np.clip(x, 1.0, float(n))
    i = np.minimum(np.floor(x).astype(np.int64), max(n - 1, 1))
    w = x - i
    lo = a[..., i - 1].astype(np.float64)
    hi = a[..., np.minimum(i, n - 1)].astype(np.float64)
    return (lo * (1 - w) + hi * w) * scale


def _range_step(t0: float, t1: float, n: int) -> float:
    """step(range(t0, t1; length = n))."""
    return (t1 - t0) / (n - 1) if n > 1 else 1.0


def load_gf_ref(f: Mapping, dom: Domain, cfg: dict, gf_of: str = "heterogeneous", tmax: float = 0.0,
                homogeneous: Optional[Mapping] = None, extrap_type: str = "pv_from_pin", dtype=np.float32,
                resample: Optional[Callable] = None, name: str = "<memory>"):
    """src/greens_io.jl:224-477 on the datasets of a reference GF file
    (and, for gf_of = "scattered", of the homogeneous companion).  `dtype`
    is Data.Array's number type.  `resample(a, t_in0, dt_in, t_out0, dt_out,
    n_out, scale)` replaces the device's cubic resample."""
    if extrap_type not in EXTRAP_TYPES:
        raise ValueError('load_gf_ref: extrap_type must be "pv_from_pin", "pv_from_pv", or "p_from_pin"; '
                         f"got {extrap_type!r}.")
    if gf_of == "scattered" and homogeneous is None:
        raise ValueError("load_gf_ref: gf_of=GFContent.scattered requires `homogeneous_path` to the "
                         "homogeneous-medium GF file.")
    s = cfg["surfaces"]
    dS_src = fib_sphere_area(s["radius_inner"], s["nPoints_inner"])
    dS_rec = fib_sphere_area(s["radius_outer"], s["nPoints_outer"])
    inj = 1.0
    k, t_raw, dom_dt, dom_dx, backend, units = _ref_datasets(f, name)
    if units != "physical":
        raise ValueError(f"load_gf_ref: {name} has units_convention={units!r}; expected \"physical\". Files "
                         "written before the units-convention change need to be regenerated (no back-compat path).")
    if gf_of == "scattered":
        kh, t_hom, _, _, _, _ = _ref_datasets(homogeneous, name)
        k = {c: k[c] - kh[c] for c in k}
        t_raw = t_hom
    n_in = t_raw.shape[0]
    dt_in = float(t_raw[1] - t_raw[0])
    t_end = float(tmax) if tmax > 0 else float(t_raw[-1])
    step_in = _range_step(float(t_raw[0]), float(t_raw[-1]), n_in)
    tout = julia_step_range(0.0, dom.dt, t_end)
    C_p, C_v, _ = scaling_factors(dom, dom_dx, dom_dt, dt_in, cfg)
    # physical -> FDTD-recorded, in Float32 (:286-289): [rec][src]
    k["p_p"] = k["p_p"] * np.float32(C_p)
    k["v_p"] = k["v_p"] * np.float32(C_p)
    k["p_v"] = k["p_v"] * np.float32(C_v)
    k["v_v"] = k["v_v"] * np.float32(C_v)
    rs = resample or _cubic
    scale = dom.dt / step_in  # step(t_out) / step(t_in)
    g = {c: np.asarray(rs(k[c], float(t_raw[0]), step_in, 0.0, dom.dt, tout.shape[0], scale)).astype(dtype)
         for c in k}
    # recording -> injection bridge (:324)
    e = (dom.c0 ** 2 * dom_dt ** 2 * dom.dt) / (C_p * dom.dx ** 3)
    # per-(emission source, recording receiver) scale: first key of gf
    # acts as the emission source at consumption (reciprocity), second as
    # the recording receiver
    sc = {"p_p": 1.0, "p_v": 1.0, "v_p": 1.0, "v_v": 1.0}
    z0 = dom.z0
    if extrap_type == "pv_from_pv":
        for a in ("p", "v"):
            sc[a + "_p"] *= z0
            sc[a + "_v"] /= z0
        for b in ("p", "v"):
            sc["p_" + b] /= z0
            sc["v_" + b] *= z0
        for c in sc:
            sc[c] *= -1
    elif extrap_type == "pv_from_pin":
        for b in ("p", "v"):
            sc["p_" + b] /= z0
            sc["v_" + b] *= z0
    for c in sc:
        sc[c] *= dS_src * dS_rec * inj * e
        g[c] = (g[c] * sc[c]).astype(dtype)  # gf .*= s in the array's type
    if extrap_type == "pv_from_pv":
        return PVFromPV(g["p_p"], g["p_v"], g["v_p"], g["v_v"])
    # the pin combinations widen to Float64 through z0 (:466-475)
    p_pin = g["p_p"].astype(np.float64) - g["p_v"].astype(np.float64) / z0
    v_pin = g["v_p"].astype(np.float64) - g["v_v"].astype(np.float64) / z0
    if extrap_type == "pv_from_pin":
        return PVFromPin(p_pin, v_pin)
    return PFromPin((p_pin + v_pin * z0) / 2, (p_pin - v_pin * z0) / 2)


def _mdd_datasets(f: Mapping, name: str):
    """`_read_gf_mdd_h5` (src/greens_io.jl:626-666): p_p and v_p (already in
    the pin form) as Float32 times Float32(empirical_scale), and t."""
    units = str(f.get("units_convention", ""))
    if "empirical_scale" not in f:
        raise ValueError(f"_read_gf_mdd_h5: {name} missing `empirical_scale` attr. "
                         "Re-extract via scripts/greens/mdd_extract.py to add it.")
    if units != "physical":
        raise ValueError(f"_read_gf_mdd_h5: {name} has units_convention={units!r}; expected \"physical\".")
    es = np.float32(float(f["empirical_scale"]))
    return (np.asarray(f["p_p"], np.float32) * es, np.asarray(f["v_p"], np.float32) * es,
            np.asarray(f["t"], np.float64))


def load_gf_mdd(f: Mapping, dom: Domain, cfg: dict, gf_of: str = "heterogeneous", tmax: float = 0.0,
                homogeneous: Optional[Mapping] = None, dtype=np.float32, name: str = "<memory>") -> PVFromPin:
    """src/greens_io.jl:536-617: MDD GFs (already incoming-pressure form)
    linearly resampled onto 0:dom.dt:t_end (tin = range(0, t[end];
    length = nt)), scaled by the two surfaces' area elements and z0 (÷ on
    the p channel, × on the v channel)."""
    if gf_of == "scattered" and homogeneous is None:
        raise ValueError("load_gf_mdd: gf_of=GFContent.scattered requires `homogeneous_path` to the "
                         "homogeneous-medium GF file.")
    s = cfg["surfaces"]
    dS = fib_sphere_area(s["radius_inner"], s["nPoints_inner"]) * fib_sphere_area(s["radius_outer"],
                                                                                    s["nPoints_outer"])
    pp, vp, t = _mdd_datasets(f, name)
    if gf_of == "scattered":
        hp, hv, t = _mdd_datasets(homogeneous, name)
        pp, vp = pp - hp, vp - hv
    n = t.shape[0]
    step_in = _range_step(0.0, float(t[-1]), n)
    t_end = float(tmax) if tmax > 0 else float(t[-1])
    tout = julia_step_range(0.0, dom.dt, t_end)
    scale = dom.dt / step_in
    gp = _linear(pp, 0.0, step_in, 0.0, dom.dt, tout.shape[0], scale).astype(dtype)
    gv = _linear(vp, 0.0, step_in, 0.0, dom.dt, tout.shape[0], scale).astype(dtype)
    e, inj = 1.0, 1.0
    sp = (1.0 / dom.z0) * (dS * inj * e)
    sv = dom.z0 * (dS * inj * e)
    return PVFromPin((gp * sp).astype(dtype), (gv * sv).astype(dtype))
