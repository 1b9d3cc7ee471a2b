/*
 * adfdtd — MI355X (gfx950) 3-D staggered-grid acoustic FDTD, C ABI.
 *
 * A drop-in device backend for the reference's `run_fdtd!` path
 * (Nano560/acoustic-disguising, Julia).  Every array argument is a HOST
 * pointer laid out exactly as the reference's Julia arrays (column-major,
 * 1-based indices where the reference stores indices).  The library copies
 * in on set_* and out on get_*, and never keeps a host pointer after a call
 * returns.  Every function returns 0 on success and a negative code on
 * error; adf_last_error() gives the message (thread-local).
 *
 * Element type of the `void*` field/trace buffers follows the solver dtype:
 * ADF_F64 -> double, ADF_F32 / ADF_F32_REF -> float.
 */
#ifndef ADFDTD_H
#define ADFDTD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ADF_ABI_VERSION 1

/* Error codes */
#define ADF_OK 0
#define ADF_EINVAL -1   /* bad argument / shape */
#define ADF_EDEVICE -2  /* HIP runtime error */
#define ADF_ENOMEM -3   /* device allocation failed or budget exceeded */
#define ADF_ESTATE -4   /* call not valid in the current state */

/* Solver element types.
 *   ADF_F64     double storage, the reference arithmetic in double
 *               (bit-reproducible against a Float64 run of the algorithm).
 *   ADF_F32     float storage and float arithmetic (fast path).
 *   ADF_F32_REF float storage with the reference's own mixed arithmetic
 *               (Float32 arrays, Float64 domain scalars: src/simulation.jl
 *               Domain fields are Float64, src/AcousticDisguising.jl:57
 *               sets Float32 arrays). */
#define ADF_F32 1
#define ADF_F64 2
#define ADF_F32_REF 3

/* The reference `Domain` (src/simulation.jl:25-44): same fields, same order,
 * so a Julia `Ref(dom)` can be passed directly. */
typedef struct adf_domain {
    double tmax, xmax, ymax, zmax;
    int64_t nx, ny, nz;
    double c0, r0, cf, dx, dy, dz, fact_m0, fact_m1, dt;
    int64_t nt;
    double z0;
} adf_domain;

/* The reference `CPMLCoefficients` (src/kernels/cpmlcoeffs.jl:12-34) of
 * `Cpml.c` (src/simulation.jl:200-227), as host doubles.  a_l..b_r have
 * npml entries, a_hl..b_hr npml+1.  Pointers may be NULL when npml == 0. */
typedef struct adf_cpml {
    int64_t npml;
    const double *a_l, *a_r, *a_hl, *a_hr, *b_l, *b_r, *b_hl, *b_hr;
} adf_cpml;

typedef struct adf_ctx adf_ctx;
typedef struct adf_solver adf_solver;
typedef struct adf_batch adf_batch;
typedef struct adf_server adf_server;

/* ---------------------------------------------------------------- context */

/* Library version (ADF_ABI_VERSION). */
int adf_version(void);
/* Message of the last failing call on this thread ("" if none). */
const char* adf_last_error(void);
/* Bind a context to HIP device `device`; owns one HIP stream. */
int adf_ctx_create(int device, adf_ctx** out);
int adf_ctx_destroy(adf_ctx* ctx);
/* Block until all work queued on the context's stream has finished. */
int adf_ctx_synchronize(adf_ctx* ctx);

/* ----------------------------------------------------------------- solver */

/* Replaces the device-side allocations of `@zeros(nx,ny,nz,4)` (the field),
 * `Cpml(dom; npml, rcoef, fc)` memory (src/simulation.jl:206-227) and the
 * CPML coefficient arrays.  Field and CPML memory start at zero. */
int adf_solver_create(adf_ctx* ctx, const adf_domain* dom, int dtype, const adf_cpml* cpml,
                      adf_solver** out);
int adf_solver_destroy(adf_solver* s);

/* field (nx,ny,nz,4): P, Vx, Vy, Vz (src/kernels/forward.jl:35-38). */
int adf_solver_set_field(adf_solver* s, const void* field);
int adf_solver_get_field(adf_solver* s, void* field);

/* `field .= 0; resetCPMLmemory!(cpml)` (src/greens.jl:161-162;
 * src/simulation.jl:235-241); also zeroes the receiver buffers. */
int adf_solver_reset(adf_solver* s);

/* CPML memory `cpml.p[a]` (extent npml+1 along axis a) and `cpml.v[a]`
 * (extent npml), each (.,.,.,2) (src/simulation.jl:213-220). */
int adf_solver_set_cpml_memory(adf_solver* s, const void* const p[3], const void* const v[3]);
int adf_solver_get_cpml_memory(adf_solver* s, void* const p[3], void* const v[3]);

/* Scatterer update mask `va` (src/geometry.jl:250-280): Bool (nx,ny,nz,4),
 * channel 1 pressure, 2..4 vx,vy,vz faces; NULL = no mask (`va === nothing`,
 * src/kernels/forward.jl:77-89,99-107). */
int adf_solver_set_mask(adf_solver* s, const uint8_t* mask4);

/* Build the same mask on the device from a shape predicate evaluated at the
 * pressure-grid coordinates (src/geometry.jl:86-143,159-176,250-280).
 * shape: 0 none, 1 sphere(radius), 2 cube(edge), 3 cross(length),
 * 4 ellipsoid(a,b,c; outside iff x^2/a^2+y^2/b^2+z^2/c^2 > 1, an addition).
 * xc/yc/zc: the n coordinates of x(d), y(d), z(d) (src/simulation.jl:89-93). */
int adf_solver_build_mask(adf_solver* s, int shape, const double* params, const double* xc,
                          const double* yc, const double* zc);
/* Copy the device mask back in the reference Bool (nx,ny,nz,4) layout. */
int adf_solver_get_mask(adf_solver* s, uint8_t* mask4);

/* `transceiversToGrid` output (src/simulation.jl:387-431) for ntx
 * transceivers, channels c = p,vx,vy,vz:
 *   idx   int64 [4][ntx][3]   ChannelBuffers.i (3,ntx), 1-based (the
 *                             reference stores them as Float32 and casts
 *                             with Int(); pass Int.(i))
 *   w     T     [4][ntx][8]   ChannelBuffers.w (2,2,2,ntx)
 *   src   T     [4][ntx][nt]  ChannelBuffers.src (nt,ntx)
 *   dir   T     [ntx][4]      TransceiverGrid.direction (4,ntx)
 *   vn_src T    [ntx][nt]     TransceiverGrid.vn_src (nt,ntx); NULL = zeros
 * Receiver buffers (rec, vn_rec) start at zero. */
int adf_solver_set_transceivers(adf_solver* s, int64_t ntx, const int64_t* idx, const void* w,
                                const void* src, const void* dir, const void* vn_src);

/* Trace buffers, each (nt,ntx) column-major in the solver dtype:
 *   which 0..3  ChannelBuffers.rec of p, vx, vy, vz
 *   which 4     vn_rec
 *   which 5..8  ChannelBuffers.src of p, vx, vy, vz (as mutated by vn2v!,
 *               src/kernels/forward.jl:130-132)
 *   which 9     vn_src
 * set_trace writes a buffer (e.g. `inject_inner_sources!`,
 * diagnostics/_check_inject_kh/fdtd_runners.jl:39-45, writes 5 and 9). */
#define ADF_REC_P 0
#define ADF_REC_VX 1
#define ADF_REC_VY 2
#define ADF_REC_VZ 3
#define ADF_VN_REC 4
#define ADF_SRC_P 5
#define ADF_SRC_VX 6
#define ADF_SRC_VY 7
#define ADF_SRC_VZ 8
#define ADF_VN_SRC 9
int adf_solver_get_trace(adf_solver* s, int which, void* out);
int adf_solver_set_trace(adf_solver* s, int which, const void* in);

/* In-loop Green's-function extrapolation (`gf_map` of run_fdtd!,
 * src/kernels/extrapolation.jl:36-249).  kind:
 *   ADF_GF_PV_FROM_PIN  tensors = {p_pin, v_pin}, each (nt_gf, nsrc, nrec)
 *   ADF_GF_PV_FROM_PV   tensors = {pp, pv, vp, vv}, each (nt_gf, nrec_in, nout)
 *   ADF_GF_P_FROM_PIN   tensors = {pos, neg}, each (nt_gf, nrec, nsrc)
 * as host doubles in Julia layout.  iRec/iSrc are 1-based transceiver
 * columns.  writeback 1 adds the extrapolated sources into the source slots
 * at it+1 (the documented intent, docs/impulsive_gf_hologram_path.md:154-171);
 * iSrc must then name distinct transceivers (ADF_EINVAL otherwise: two
 * outputs would write the same source slots).  0 leaves the source slots as
 * they are: the reference computes the sources into copies and discards
 * them (extrapolation.jl:215-216,226-232,246-247), which has no observable
 * result, so nothing is computed.  kind 0 clears the map. */
#define ADF_GF_NONE 0
#define ADF_GF_PV_FROM_PIN 1
#define ADF_GF_PV_FROM_PV 2
#define ADF_GF_P_FROM_PIN 3
int adf_solver_set_gf_map(adf_solver* s, int kind, const double* const* tensors, int64_t nt_gf,
                          int64_t dim2, int64_t dim3, const int64_t* iRec, int64_t n_irec,
                          const int64_t* iSrc, int64_t n_isrc, int writeback);

/* `for it in start_step:end_step; forward_onestep!(...); end`
 * (src/simulation.jl:542-557; src/kernels/forward.jl:13-139).
 * Steps are 1-based as in the reference.  Synchronous on return. */
int adf_solver_run(adf_solver* s, int64_t start_step, int64_t end_step);

/* One pressure slice (the `on_step` snapshots of src/hologram.jl:238-247):
 * axis 0 -> P[index,:,:] (ny,nz); 1 -> P[:,index,:] (nx,nz);
 * 2 -> P[:,:,index] (nx,ny); index is 1-based. */
int adf_solver_get_pressure_slice(adf_solver* s, int axis, int64_t index, void* out);

/* Kernel path and tiling (tuning knobs; results do not depend on them):
 *   ADF_OPT_PATH     0 auto (fused when a field buffer is under 4 GiB and a
 *                    row needs at most 63 segments of 248 cells, i.e.
 *                    nx <= 15624), 1 two-pass P/V kernels, 2 fused
 *                    z-marching kernel
 *   ADF_OPT_TILE_Y   rows per wave of the fused kernel: 0 auto, 2 (the only
 *                    shape built)
 *   ADF_OPT_BLOCK_Z  planes per workgroup of the fused kernel: 0 auto, >= 1 */
#define ADF_OPT_PATH 1
#define ADF_OPT_TILE_Y 2
#define ADF_OPT_BLOCK_Z 3
/*   ADF_OPT_TIME_SWEEP  v >= 1: record HIP events around the field-update
 *                        launches (fused sweep, or the P and V kernels of
 *                        the two-pass path) of every v-th step of a run
 *                        (the first step of the run included) for
 *                        adf_solver_sweep_stats; 0: off (default).  Each
 *                        recorded event pair adds device time to its step
 *                        (about 18 us at 256^3), so sampled timing (v of
 *                        16 or more) keeps the run's own step time. */
#define ADF_OPT_TIME_SWEEP 4
/*   ADF_OPT_WAVES    waves per workgroup of the fused kernel: 0 auto, 1, 2, 4
 *                    (a workgroup takes that many y-neighbouring tiles) */
#define ADF_OPT_WAVES 5
/*   ADF_OPT_BLOCK_Z_GEN  planes per tile of the fused kernel's side tiles
 *                        (CPML slabs in y/z, boundary rows and planes):
 *                        0 auto, >= 1 */
#define ADF_OPT_BLOCK_Z_GEN 6
/*   ADF_OPT_PIPELINE     1 (default): the fused kernel's regular tiles load
 *                        the next plane while the current one is computed
 *                        (two plane buffers in registers); 0: one plane at
 *                        a time */
#define ADF_OPT_PIPELINE 7
/*   ADF_OPT_SWEEP_LAUNCH how the fused kernel's tiles are launched:
 *                        ADF_SWEEP_ONE_LAUNCH (default) one launch over all
 *                        tiles; ADF_SWEEP_TWO_STREAMS side tiles (CPML slabs,
 *                        boundary) and interior tiles as two launches on two
 *                        streams at the same time; ADF_SWEEP_TWO_LAUNCHES
 *                        the same two launches one after the other */
#define ADF_OPT_SWEEP_LAUNCH 8
#define ADF_SWEEP_ONE_LAUNCH 0
#define ADF_SWEEP_TWO_STREAMS 1
#define ADF_SWEEP_TWO_LAUNCHES 2
/*   ADF_OPT_STORE_NT     1: the fused kernel's pipelined routines store
 *                        field rows with the non-temporal hint (faster
 *                        where the two field buffers exceed the GPU's
 *                        last-level cache, e.g. 256^3; slower where they
 *                        fit); 0: ordinary stores; -1 (default):
 *                        automatic, i.e. ordinary stores for a solve run
 *                        alone, and in a batch the hint when the members'
 *                        field buffers together exceed 256 MB.  A batch
 *                        takes each member's setting.  Results are
 *                        identical. */
#define ADF_OPT_STORE_NT 9
/*   ADF_OPT_TX_LAUNCH    1: on the fused path, the transceivers of a step
 *                        run in one launch when the ones that inject or
 *                        read injected cells are few (at most 8 levels and
 *                        512 entries per phase, 4096 fix-up cells): one
 *                        workgroup runs those in order, the others record;
 *                        2 (default): as 1, but inside the sweep launch
 *                        (one launch, one wave per workgroup, no GF map, at
 *                        most 64 tiles holding cells those transceivers
 *                        touch; otherwise as 1): their workgroups wait for
 *                        the tiles that store those cells, and the others'
 *                        records are taken by the next step's sweep launch
 *                        (the last step's when the run returns); 0: one
 *                        launch per transceiver level and one for the
 *                        velocity fix-up.  Results are identical. */
#define ADF_OPT_TX_LAUNCH 10
/*   ADF_OPT_CELLS_PER_LANE 0 (default): the fused kernel's lanes hold 1
 *                        cell up to nx = 64, 2 up to 128, 4 above; or 1, 2
 *                        or 4.  Rows wider than 64 lanes are cut into
 *                        segments of at most 62 lanes, one wave each.
 *                        Results are identical. */
#define ADF_OPT_CELLS_PER_LANE 11
int adf_solver_set_option(adf_solver* s, int option, int64_t value);

/* Mean device time (ms) of one field-update launch and the number of
 * launches timed during the last adf_solver_run with ADF_OPT_TIME_SWEEP
 * on (measurement support; no counterpart in the reference). */
int adf_solver_sweep_stats(adf_solver* s, double* avg_ms, int64_t* launches);

/* Mean device time per step of the last adf_solver_run (ms), measured
 * with HIP events on the solver's stream. */
int adf_solver_last_step_ms(adf_solver* s, double* ms);

/* The step the solver takes next (1-based): 1 after create/reset, end+1
 * after adf_solver_run(start, end), advanced by adf_batch_run. */
int adf_solver_next_step(adf_solver* s, int64_t* it);

/* ------------------------------------------------------------------ batch
 *
 * Many independent solves advanced together on one context.  An addition:
 * the reference runs its solves one after another (impulsive_gfs loops
 * run_fdtd! over 2*n_src sources, src/greens.jl:143-164; build_reverb_data
 * over its sources, src/reverb.jl:265-312).  Each member keeps its own
 * domain, CPML, mask, transceivers and step counter; a batched step of a
 * member is the same forward_onestep! (src/kernels/forward.jl:13-139) that
 * adf_solver_run takes, value for value, whatever else is in the batch.
 * One step of the batch runs the field sweep of all members as two launches
 * (interior tiles, side tiles) per row-width class, the transceiver levels
 * of all members as one launch per level, then the velocity fixups and
 * mirrors; it is captured as a HIP graph and replayed.
 * Members: same context and dtype, a fused sweep (4 nx ny nz sizeof(T) <
 * 4 GiB, nx <= 15624), no GF map. */
int adf_batch_create(adf_ctx* ctx, adf_batch** out);
/* Members are released (not destroyed). */
int adf_batch_destroy(adf_batch* b);
int adf_batch_add(adf_batch* b, adf_solver* s);
int adf_batch_remove(adf_batch* b, adf_solver* s);
int adf_batch_size(adf_batch* b, int64_t* n);
/* Advance every member by nsteps steps from its own next step
 * (adf_solver_next_step); fails if a member would pass its nt.
 * Synchronous on return. */
int adf_batch_run(adf_batch* b, int64_t nsteps);
/*   ADF_BATCH_PERSISTENT   1: a fixed grid of waves drains an atomic tile
 *                          queue; 0 (default): one workgroup per tile
 *   ADF_BATCH_GRAPH_STEPS  steps per captured HIP graph (default 16; 0 runs
 *                          the launches directly) */
#define ADF_BATCH_PERSISTENT 1
#define ADF_BATCH_GRAPH_STEPS 2
int adf_batch_set_option(adf_batch* b, int option, int64_t value);
/* Mean device time per batched step of the last adf_batch_run (ms). */
int adf_batch_last_step_ms(adf_batch* b, double* ms);

/* ------------------------------------------- impulsive_gfs post-processing
 *
 * The per-run tail of impulsive_gfs (src/greens.jl:166-191) on the device,
 * for a solver whose run has finished: the p and vn receiver traces of
 * transceivers first .. first+n_rec-1 (0-based; first = 1 drops the
 * source's self-record as greens.jl:171-174 does), denoised (a sample is
 * zeroed where |A| < 1e-4 * the trace's signed maximum over time,
 * greens.jl:176-177) and resampled (resample_time(A, t_shifted, t_gf;
 * kind = :cubic), greens.jl:183, src/resample.jl:35-77: the cubic B-spline
 * with Line(OnGrid()) ends through A on t_shifted = range(-timeshift,
 * tmax - timeshift; length = nt), Flat outside, sampled on t_gf =
 * range(0, tmax; length = nt_gf) and scaled by step(t_gf)/step(t_shifted)).
 * out_p / out_vn: float32 (nt_gf, n_rec) column-major each, the layout of
 * the reference's p_p[:, :, iSrc] / v_p[:, :, iSrc] slices. */
int adf_solver_gf_resample(adf_solver* s, int64_t first, int64_t n_rec, double timeshift, int64_t nt_gf,
                           float* out_p, float* out_vn);

/* Cubic resampling of host float32 traces on the device, the reverb
 * stage's per-source resample (src/reverb.jl:424-452 _resample_iSrc!,
 * src/resample.jl:35-77 resample_time(kind = :cubic)): in holds ncol
 * traces of n_in samples (column-major (n_in, ncol)) on t_in = t_in0 +
 * (k-1) dt_in; each is interpolated by the natural cubic B-spline
 * (Interpolations' BSpline(Cubic(Line(OnGrid())))), Flat outside, sampled
 * at t_out = t_out0 + (k-1) dt_out, k = 1..n_out, multiplied by amp_scale
 * (1 for scale_amplitude = false) and rounded to float32 into out
 * (column-major (n_out, ncol)).  denoise = 1 first zeroes the samples below
 * 1e-4 of each trace's signed maximum (src/greens.jl:176-177).  The spline
 * is solved and evaluated in double.  Synchronous on ctx's stream. */
int adf_resample_f32(adf_ctx* ctx, const float* in, int64_t n_in, int64_t ncol, double t_in0, double dt_in,
                     double t_out0, double dt_out, int64_t n_out, double amp_scale, int denoise, float* out);

/* adf_resample_f32 for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  in and out are DEVICE
 * pointers the caller owns (same layouts as adf_resample_f32); the pivot
 * sequence and the spline coefficients live in a fixed 32 MiB device array
 * of the library (one per device, held until the process ends), so
 * (n_in - 2 rounded up to 32) + ncol * (n_in + 2) must not exceed 4 Mi
 * doubles (ADF_EINVAL otherwise).  Calls no hipMalloc/hipFree and
 * allocates no host memory; enqueues three kernels on ctx's stream and
 * returns: out is complete once the stream reaches that point
 * (adf_ctx_synchronize).  Every call on one device shares the array:
 * calls that overlap on one device (two contexts, or two threads) race. */
int adf_resample_f32_noalloc(adf_ctx* ctx, const float* in, int64_t n_in, int64_t ncol, double t_in0,
                             double dt_in, double t_out0, double dt_out, int64_t n_out, double amp_scale,
                             int denoise, float* out);

/* --------------------------------------------------------- post-hoc KH
 *
 * Analytic Kirchhoff-Helmholtz kernels evaluated on the device and
 * convolved with surface traces, replacing eval_one_kernel
 * (src/kh_kernels.jl:106-157) + eval_kernel_pedestal (:174-203) +
 * direct_kh_accumulate! (:256-296) without forming the dense kernel:
 *   out[t,s] += alpha * sum_r sum_{tt=1}^{min(t,ntK)} K(tt;r,s) in[t-tt+1,r]
 * which: 0 p_p, 1 p_v, 2 v_p, 3 v_v, at t_grid[tt] (ntK increasing times)
 * with Gaussian width sigma_t, sound speed c0, density rho0.  pedestal = 1
 * (v_v): K - K_dc plus alpha K_dc[r,s] cumsum(in)[t,r], as the reference's
 * truncation-friendly form.  Points and normals are the Julia (n,3)
 * matrices (column-major); in (nt, n_outer) and out (nt, n_inner) are
 * column-major host doubles, out accumulated.  dtype ADF_F64: kernel values
 * and products in double; ADF_F32 / ADF_F32_REF: kernel values rounded to
 * Float32 as eval_one_kernel(; T = Float32), inputs rounded to Float32,
 * Float32 products summed in double.  Synchronous. */
int adf_kh_analytic_accumulate(adf_ctx* ctx, int which, int dtype, int64_t n_outer, const double* outer_pts,
                               const double* outer_nrm, int64_t n_inner, const double* inner_pts,
                               const double* inner_nrm, const double* t_grid, int64_t ntK, double sigma_t,
                               double c0, double rho0, int pedestal, double alpha, int64_t nt, const double* in,
                               double* out);
/* Device time (ms) of the prefix-sum and convolution kernels of this
 * thread's last adf_kh_analytic_accumulate (measurement support). */
int adf_kh_last_ms(double* ms);

/* ------------------------------------------------- bandwidth probes
 *
 * Measurement support (no counterpart in the reference): streaming kernels
 * over two device buffers of `bytes` each (a multiple of 16 KiB, at most
 * 1 GiB) with the 128-bit access instructions of the field-update sweep;
 * *ms_per_launch is the mean device time of `reps` launches after one
 * warm-up launch.  A copy moves 2 * bytes per launch, read and write
 * kinds `bytes`. */
#define ADF_BW_COPY_BUF 0    /* copy, buffer loads and stores */
#define ADF_BW_COPY_GLOBAL 1 /* copy, global loads and stores */
#define ADF_BW_READ 2        /* read only */
#define ADF_BW_WRITE 3       /* write only */
#define ADF_BW_COPY_NT 4     /* copy, non-temporal stores */
#define ADF_BW_ZMARCH 5      /* the fused sweep's access pattern without its arithmetic: four channels of
                                256-cell rows, 2-row x 4-plane tiles marched in z with halo rows, one plane
                                prefetched; bytes (a multiple of 4 MiB) per buffer, launches alternate */
#define ADF_BW_ZMARCH_NT 6   /* as ADF_BW_ZMARCH with non-temporal stores */
int adf_bandwidth_probe(adf_ctx* ctx, int kind, int64_t bytes, int reps, double* ms_per_launch);

/* ----------------------------------------------------------------- server
 *
 * Served form (an addition): a request loop on its own host thread and HIP
 * context.  Requests (one run_fdtd! each: src/simulation.jl:542-557, with
 * the set-up of src/greens.jl:151-164 or src/hologram.jl:215-248) are
 * admitted in arrival order while their device footprint fits the budget,
 * join the batch of their dtype, advance together with every other live
 * request in chunks of steps, and are retired when they reach end_step:
 * the requested traces are copied to host memory and the device memory is
 * released.  A request's results are the same, value for value, as an
 * adf_solver_run of the same inputs, whatever else is in flight.  No GF map
 * (those solves run alone through adf_solver_run).  Latency is measured
 * from submission to retirement. */
typedef struct adf_request {
    adf_domain dom;
    int dtype;
    adf_cpml cpml;
    /* scatterer: 0 none, 1..4 shapes of adf_solver_build_mask (params, and
     * the n coordinates xc/yc/zc of x(d), y(d), z(d)), -1 explicit Bool
     * (nx,ny,nz,4) mask4 */
    int mask_shape;
    double mask_params[4];
    const double* xc;
    const double* yc;
    const double* zc;
    const uint8_t* mask4;
    /* transceivers as adf_solver_set_transceivers */
    int64_t ntx;
    const int64_t* idx;
    const void* w;
    const void* src;
    const void* dir;
    const void* vn_src;
    /* initial field (nx,ny,nz,4), NULL = zeros (src/simulation.jl:468-518
     * initialField for holograms) */
    const void* field;
    int64_t start_step, end_step;
    /* bitmask of trace buffers to return (bit b = buffer id b of
     * adf_solver_get_trace); 0 = p, vx, vy, vz and vn receivers (0x1f) */
    uint32_t want;
    int keep_field; /* 1: also return the final field */
} adf_request;

#define ADF_REQ_QUEUED 0
#define ADF_REQ_RUNNING 1
#define ADF_REQ_DONE 2
#define ADF_REQ_FAILED -1

typedef struct adf_server_info {
    int64_t completed, failed, running, queued;
    uint64_t bytes_in_use, peak_bytes, budget_bytes;
    double p50_ms, p99_ms, mean_ms; /* submission to retirement */
    double solves_per_s;            /* completed / (last completion - first submission) */
    double busy_s;                  /* time spent stepping */
    double glups;                   /* cell updates per busy second / 1e9 */
} adf_server_info;

/* hbm_budget_bytes: device bytes live requests may hold (0 = 85 % of the
 * free memory at creation). */
int adf_server_create(int device, uint64_t hbm_budget_bytes, adf_server** out);
/* Finishes running requests, fails queued ones, releases everything. */
int adf_server_destroy(adf_server* sv);
/*   ADF_SERVER_CHUNK   steps per scheduling round (default 64)
 *   ADF_SERVER_BUDGET  device-memory budget in bytes */
#define ADF_SERVER_CHUNK 1
#define ADF_SERVER_BUDGET 2
int adf_server_set_option(adf_server* sv, int option, int64_t value);
/* Copies every input; returns at once with the request id. */
int adf_server_submit(adf_server* sv, const adf_request* req, int64_t* id);
int adf_server_poll(adf_server* sv, int64_t id, int* status);
/* Block until done or failed, or timeout_s (< 0: no timeout). */
int adf_server_wait(adf_server* sv, int64_t id, double timeout_s, int* status);
/* Results of a done request: trace buffer `which` ((nt,ntx) as
 * adf_solver_get_trace) and, with keep_field, the final field. */
int adf_server_get_trace(adf_server* sv, int64_t id, int which, void* out);
int adf_server_get_field(adf_server* sv, int64_t id, void* out);
/* Message of a failed request ("" otherwise); valid until release. */
const char* adf_server_request_error(adf_server* sv, int64_t id);
/* Forget a done or failed request and free its host results. */
int adf_server_release(adf_server* sv, int64_t id);
int adf_server_stats(adf_server* sv, adf_server_info* st);

#ifdef __cplusplus
}
#endif

#endif /* ADFDTD_H */
