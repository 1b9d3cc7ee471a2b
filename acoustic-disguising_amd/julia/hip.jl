# HIP backend of the reference's FDTD path over libadfdtd.so (include/adfdtd.h).
#
# Installed into the reference as src/backends/hip.jl and selected the way
# the existing backends are, by ACOUSTIC_DISGUISING_BACKEND at module load
# (src/AcousticDisguising.jl:50-61):
#
#     elseif BACKEND == "HIP"
#         @init_parallel_stencil(Threads, Float32, 2, inbounds = false)
#         const DeviceArray = Array
#         include("backends/hip.jl")
#
# and in src/simulation.jl the existing run_fdtd! (:542-557) forwards to it:
#
#     @static if BACKEND == "HIP"
#         run_fdtd!(dom, field, txs, cpml, va; kw...) = HIPBackend.run_fdtd!(dom, field, txs, cpml, va; kw...)
#     else
#         ... the existing method ...
#     end
#
# Arrays stay host `Array`s (Data.Array under the Threads backend); the library
# copies them to the device on the way in and back on the way out and keeps no
# host pointer.  Public types and signatures of the reference are unchanged.
# ENV["ADFDTD_LIB"] names the library (default: libadfdtd.so on the loader path).

module HIPBackend

const LIB = get(ENV, "ADFDTD_LIB", "libadfdtd.so")
const ADF_F32, ADF_F64, ADF_F32_REF = Cint(1), Cint(2), Cint(3)
# trace buffer ids of adf_solver_get_trace / set_trace
const REC_P, VN_REC, SRC_P, VN_SRC = Cint(0), Cint(4), Cint(5), Cint(9)

# adf_cpml: npml and the eight coefficient vectors of Cpml.c (Float64)
struct AdfCpml
    npml::Int64
    a_l::Ptr{Float64}; a_r::Ptr{Float64}; a_hl::Ptr{Float64}; a_hr::Ptr{Float64}
    b_l::Ptr{Float64}; b_r::Ptr{Float64}; b_hl::Ptr{Float64}; b_hr::Ptr{Float64}
end

check(rc) = rc == 0 || error("libadfdtd: ", unsafe_string(ccall((:adf_last_error, LIB), Cstring, ())))

const CTX = Ref{Ptr{Cvoid}}(C_NULL)
"""The library context (device 0 and its stream), created on first use."""
function ctx()
    if CTX[] == C_NULL
        r = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:adf_ctx_create, LIB), Cint, (Cint, Ref{Ptr{Cvoid}}), 0, r))
        CTX[] = r[]
    end
    return CTX[]
end

"""Solver element type for arrays of type T: Float64 runs the Float64
algorithm; Float32 the reference's own Float32 arithmetic, or the fast path
with `fast = true`."""
dtype_code(::Type{Float64}; fast::Bool = false) = ADF_F64
dtype_code(::Type{Float32}; fast::Bool = false) = fast ? ADF_F32 : ADF_F32_REF

# ------------------------------------------------------------------ solver

"""adf_solver_create: device field, CPML memory and coefficients of one
solve.  `Domain` is isbits with the field order of adf_domain."""
function create(dom, cpml, code::Cint)
    c = cpml.c
    arrs = map(v -> Vector{Float64}(v), (c.a_l, c.a_r, c.a_hl, c.a_hr, c.b_l, c.b_r, c.b_hl, c.b_hr))
    cp = AdfCpml(Int64(size(cpml.v[1], 1)), map(pointer, arrs)...)
    s = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve arrs check(ccall((:adf_solver_create, LIB), Cint,
        (Ptr{Cvoid}, Ref{typeof(dom)}, Cint, Ref{AdfCpml}, Ref{Ptr{Cvoid}}), ctx(), dom, code, cp, s))
    return s[]
end

destroy(s) = ccall((:adf_solver_destroy, LIB), Cint, (Ptr{Cvoid},), s)

set_field!(s, field::Array{T,4}) where {T} =
    check(ccall((:adf_solver_set_field, LIB), Cint, (Ptr{Cvoid}, Ptr{T}), s, field))
get_field!(field::Array{T,4}, s) where {T} =
    check(ccall((:adf_solver_get_field, LIB), Cint, (Ptr{Cvoid}, Ptr{T}), s, field))

function set_cpml!(s, cpml)
    p = Ptr{Cvoid}[pointer(a) for a in cpml.p]
    v = Ptr{Cvoid}[pointer(a) for a in cpml.v]
    GC.@preserve cpml check(ccall((:adf_solver_set_cpml_memory, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}), s, p, v))
end

function get_cpml!(cpml, s)
    p = Ptr{Cvoid}[pointer(a) for a in cpml.p]
    v = Ptr{Cvoid}[pointer(a) for a in cpml.v]
    GC.@preserve cpml check(ccall((:adf_solver_get_cpml_memory, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}), s, p, v))
end

"""The Bool (nx,ny,nz,4) update mask of build_update_mask (geometry.jl:250-280),
or nothing for the unmasked path."""
function set_mask!(s, va)
    if va === nothing
        return check(ccall((:adf_solver_set_mask, LIB), Cint, (Ptr{Cvoid}, Ptr{UInt8}), s, C_NULL))
    end
    m = eltype(va) == Bool ? reinterpret(UInt8, va) : UInt8.(va .!= 0)
    GC.@preserve m check(ccall((:adf_solver_set_mask, LIB), Cint, (Ptr{Cvoid}, Ptr{UInt8}), s, m))
end

"""The packed TransceiverGrid (simulation.jl:387-431): indices (3,ntx) per
channel (stored as T by the reference), weights (2,2,2,ntx), sources
(nt,ntx), directions (4,ntx), vn sources (nt,ntx)."""
function upload_transceivers!(s, txs, ::Type{T}) where {T}
    chans = (txs.p, txs.vx, txs.vy, txs.vz)
    ntx = size(txs.direction, 2)
    idx = Int64.(cat((c.i for c in chans)...; dims = 3))                  # (3, ntx, 4), 1-based
    w = T.(cat((reshape(c.w, 8, :) for c in chans)...; dims = 3))         # (8, ntx, 4)
    src = T.(cat((c.src for c in chans)...; dims = 3))                    # (nt, ntx, 4)
    dir = T.(txs.direction)                                               # (4, ntx)
    vns = T.(txs.vn_src)                                                  # (nt, ntx)
    check(ccall((:adf_solver_set_transceivers, LIB), Cint,
        (Ptr{Cvoid}, Int64, Ptr{Int64}, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{T}), s, ntx, idx, w, src, dir, vns))
end

"""Records, vn records and the source buffers as vn2v! left them
(forward.jl:130-132) back into txs."""
function download_transceivers!(txs, s, ::Type{T}) where {T}
    buf = Array{T}(undef, size(txs.vn_rec))
    for (k, c) in enumerate((txs.p, txs.vx, txs.vy, txs.vz))
        check(ccall((:adf_solver_get_trace, LIB), Cint, (Ptr{Cvoid}, Cint, Ptr{T}), s, REC_P + k - 1, buf))
        c.rec .= buf
        check(ccall((:adf_solver_get_trace, LIB), Cint, (Ptr{Cvoid}, Cint, Ptr{T}), s, SRC_P + k - 1, buf))
        c.src .= buf
    end
    check(ccall((:adf_solver_get_trace, LIB), Cint, (Ptr{Cvoid}, Cint, Ptr{T}), s, VN_REC, buf))
    txs.vn_rec .= buf
    check(ccall((:adf_solver_get_trace, LIB), Cint, (Ptr{Cvoid}, Cint, Ptr{T}), s, VN_SRC, buf))
    txs.vn_src .= buf
end

"""The gf_map keyword (extrapolation.jl:36-249).  writeback = 1 adds the
extrapolated sources at it+1 as documented; 0 computes and discards them
as the reference's copying views do."""
function set_gf_map!(s, gf; writeback::Integer = 1)
    if gf === nothing
        return check(ccall((:adf_solver_set_gf_map, LIB), Cint,
            (Ptr{Cvoid}, Cint, Ptr{Ptr{Float64}}, Int64, Int64, Int64, Ptr{Int64}, Int64, Ptr{Int64}, Int64, Cint),
            s, 0, C_NULL, 0, 0, 0, C_NULL, 0, C_NULL, 0, 1))
    end
    kind, ts = hasproperty(gf, :p_pin) ? (1, (gf.p_pin, gf.v_pin)) :
               hasproperty(gf, :pp) ? (2, (gf.pp, gf.pv, gf.vp, gf.vv)) : (3, (gf.pos, gf.neg))
    t64 = map(t -> Array{Float64}(t), ts)
    ng, d2, d3 = size(t64[1])
    ptrs = Ptr{Float64}[pointer(t) for t in t64]
    iRec, iSrc = Int64.(gf.iRec), Int64.(gf.iSrc)
    GC.@preserve t64 check(ccall((:adf_solver_set_gf_map, LIB), Cint,
        (Ptr{Cvoid}, Cint, Ptr{Ptr{Float64}}, Int64, Int64, Int64, Ptr{Int64}, Int64, Ptr{Int64}, Int64, Cint),
        s, kind, ptrs, ng, d2, d3, iRec, length(iRec), iSrc, length(iSrc), writeback))
end

run!(s, first::Integer, last::Integer) =
    check(ccall((:adf_solver_run, LIB), Cint, (Ptr{Cvoid}, Int64, Int64), s, first, last))

"""Pressure on the plane index (1-based) of axis 1 (x: (ny,nz)), 2 (y:
(nx,nz)) or 3 (z: (nx,ny)), build_hologram's on_step slices
(hologram.jl:238-247) without pulling the field."""
function pressure_slice(s, dom, axis::Integer, index::Integer, ::Type{T}) where {T}
    shp = axis == 1 ? (dom.ny, dom.nz) : axis == 2 ? (dom.nx, dom.nz) : (dom.nx, dom.ny)
    out = Matrix{T}(undef, shp...)
    check(ccall((:adf_solver_get_pressure_slice, LIB), Cint, (Ptr{Cvoid}, Cint, Int64, Ptr{T}),
                s, axis - 1, index, out))
    return out
end

# ---------------------------------------------------------------- run_fdtd!

"""run_fdtd! with the reference's signature (simulation.jl:542-557).
on_step(it, field) gets the whole field after every step, as the reference's
callback does; `slices = (every, ix, iy, iz, push!)` instead calls
push!(it, xs, ys, zs) with the pressure planes every `every` steps (the
build_hologram snapshot) without pulling the field."""
function run_fdtd!(dom, field::Array{T,4}, txs, cpml, va; start_step::Int = 1, on_step = nothing,
                   gf_map = nothing, fast::Bool = false, slices = nothing) where {T}
    s = create(dom, cpml, dtype_code(T; fast = fast))
    try
        set_field!(s, field)
        set_cpml!(s, cpml)
        set_mask!(s, va)
        upload_transceivers!(s, txs, T)
        set_gf_map!(s, gf_map)
        if on_step !== nothing
            for it in start_step:dom.nt
                run!(s, it, it)
                get_field!(field, s)
                on_step(it, field)
            end
        elseif slices !== nothing
            every, ix, iy, iz, sink = slices
            it = start_step
            while it <= dom.nt
                last = min(dom.nt, cld(it, every) * every)
                run!(s, it, last)
                if last % every == 0
                    sink(last, [pressure_slice(s, dom, 1, i, T) for i in ix],
                         [pressure_slice(s, dom, 2, i, T) for i in iy], [pressure_slice(s, dom, 3, i, T) for i in iz])
                end
                it = last + 1
            end
        else
            run!(s, start_step, dom.nt)
        end
        get_field!(field, s)
        get_cpml!(cpml, s)
        download_transceivers!(txs, s, T)
    finally
        destroy(s)
    end
    return field
end

# ------------------------------------------------------------------ batches

"""Many solves advanced together (adf_batch_*): e.g. the 2 n_src runs of
impulsive_gfs (greens.jl:143-164).  Each member's results equal its run
through run_fdtd!."""
function batch_create()
    b = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:adf_batch_create, LIB), Cint, (Ptr{Cvoid}, Ref{Ptr{Cvoid}}), ctx(), b))
    return b[]
end
batch_add!(b, s) = check(ccall((:adf_batch_add, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), b, s))
batch_run!(b, nsteps::Integer) = check(ccall((:adf_batch_run, LIB), Cint, (Ptr{Cvoid}, Int64), b, nsteps))
batch_destroy(b) = ccall((:adf_batch_destroy, LIB), Cint, (Ptr{Cvoid},), b)

"""The per-run tail of impulsive_gfs (greens.jl:166-191) on the device for
a finished solve whose transceivers are [source, receivers...]: self-record
dropped, denoised, resampled onto nt_gf samples; returns (p, vn) Float32
(nt_gf, n_rec)."""
function gf_resample(s, n_rec::Integer, timeshift::Real, nt_gf::Integer)
    gp = Array{Float32}(undef, nt_gf, n_rec)
    gv = similar(gp)
    check(ccall((:adf_solver_gf_resample, LIB), Cint,
                (Ptr{Cvoid}, Int64, Int64, Float64, Int64, Ptr{Float32}, Ptr{Float32}),
                s, 1, n_rec, timeshift, nt_gf, gp, gv))
    return gp, gv
end

"""resample_time(A, tin, tout; kind = :cubic, scale_amplitude) (resample.jl:
35-77) of Float32 traces A (nt_in, ncol) on the device; tin = t_in0 + k dt_in,
tout = t_out0 + k dt_out, output times amp."""
function resample_f32(A::Matrix{Float32}, t_in0, dt_in, t_out0, dt_out, n_out::Integer; amp = 1.0,
                      denoise::Bool = false)
    R = Matrix{Float32}(undef, n_out, size(A, 2))
    check(ccall((:adf_resample_f32, LIB), Cint,
                (Ptr{Cvoid}, Ptr{Float32}, Int64, Int64, Float64, Float64, Float64, Float64, Int64, Float64, Cint,
                 Ptr{Float32}),
                ctx(), A, size(A, 1), size(A, 2), t_in0, dt_in, t_out0, dt_out, n_out, amp, denoise, R))
    return R
end

end # module
