// Cubic resampling of recorded traces on the device:
//
// impulsive_gfs post-processing of one finished run (include/adfdtd.h
// adf_solver_gf_resample; src/greens.jl:166-191):
//   A      = receiver traces p and vn of transceivers first .. first+n_rec-1
//            (greens.jl:171-174; first = 1 drops the source's self-record)
//   A     := 0 where |A| < 1e-4 max_t A, the signed maximum of each trace
//            (greens.jl:176-177)
//   A_gf   = resample_time(A, t_shifted, t_gf; kind = :cubic)
//            (greens.jl:183; src/resample.jl:35-77): the cubic B-spline with
//            Line(OnGrid()) ends (the natural spline) through A on
//            t_shifted = range(-timeshift, tmax - timeshift; length = nt),
//            Flat outside, sampled on t_gf = range(0, tmax; length =
//            nt_gf), times step(t_gf) / step(t_shifted), stored as float32.
//
// the reverb stage's per-source resample (adf_resample_f32;
// src/reverb.jl:424-452): the same interpolant through float32 traces the
// host has already cast and scaled, without denoise or amplitude scaling.
//
// One thread per trace solves the spline system along time (Thomas
// algorithm in double, the shared pivot sequence computed on the host);
// one thread per output sample evaluates the four B-spline taps.  Along the
// receiver and channel axes the reference's 3-D interpolant is evaluated
// only at integer nodes, where the interpolant equals the data.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/adfdtd.h"
#include "adf_internal.hpp"

using namespace adf;

namespace {

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

// Solve for the spline coefficients c_0 .. c_{n+1} of column col (rows
// (c_{i-1} + 4 c_i + c_{i+1}) / 6 = y_i, c_1 = y_1, c_n = y_n, zero second
// difference at both ends), after the denoise when asked.  Columns
// 0 .. nsplit-1 are traces first .. of a0, the rest traces first .. of a1
// (each n samples long).  cp: pivots 1/(4 - cp_{i-1}), cp_0 = 1/4.
template <typename T>
__global__ void k_rs_coef(const T* __restrict__ a0, const T* __restrict__ a1, int64_t first, int nsplit, int n,
                          int ncol, int denoise, const double* __restrict__ cp, double* __restrict__ coef) {
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= ncol) return;
    const T* a = col < nsplit ? a0 + (size_t)(first + col) * n : a1 + (size_t)(first + col - nsplit) * n;
    double* c = coef + (size_t)col * (n + 2);
    double thr = -1.0;  // no sample is below it
    if (denoise) {
        double mx = (double)a[0];
        for (int k = 1; k < n; ++k) mx = fmax(mx, (double)a[k]);
        thr = 1e-4 * mx;
    }
    auto y = [&](int k) {  // (denoised) sample k (0-based)
        const double v = (double)a[k];
        return fabs(v) < thr ? 0.0 : v;
    };
    if (n == 1) {
        c[0] = c[1] = c[2] = y(0);
        return;
    }
    c[1] = y(0);
    c[n] = y(n - 1);
    const int m = n - 2;  // unknowns c_2 .. c_{n-1}, forward sweep into c
    for (int i = 0; i < m; ++i) {
        double rhs = 6.0 * y(i + 1);
        if (i == 0) rhs -= c[1];
        if (i == m - 1) rhs -= c[n];
        c[2 + i] = i == 0 ? rhs / 4.0 : (rhs - c[1 + i]) / (4.0 - cp[i - 1]);
    }
    for (int i = m - 2; i >= 0; --i) c[2 + i] = c[2 + i] - cp[i] * c[3 + i];
    c[0] = 2.0 * c[1] - c[2];
    c[n + 1] = 2.0 * c[n] - c[n - 1];
}

// out[col][k] = scale * s_col(x_k), x_k = 1 + (t_out0 + k dt_out - t_in0) / dt_in
// clamped to [1, n].
__global__ void k_rs_eval(const double* __restrict__ coef, int n, int ncol, double t_in0, double dt_in,
                          double t_out0, double dt_out, int n_out, double scale, float* __restrict__ out) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (int64_t)ncol * n_out) return;
    const int col = (int)(g / n_out), k = (int)(g - (int64_t)col * n_out);
    const double* c = coef + (size_t)col * (n + 2);
    double x = 1.0 + ((t_out0 + (double)k * dt_out) - t_in0) / dt_in;
    x = fmin(fmax(x, 1.0), (double)n);
    int i = (int)floor(x);
    if (i > n - 1) i = n - 1 > 1 ? n - 1 : 1;
    const double d = x - (double)i, e = 1.0 - d;
    const double w0 = e * e * e / 6.0;
    const double w1 = 2.0 / 3.0 - d * d + d * d * d / 2.0;
    const double w2 = 2.0 / 3.0 - e * e + e * e * e / 2.0;
    const double w3 = d * d * d / 6.0;
    auto cc = [&](int q) { return c[q < 0 ? 0 : (q > n + 1 ? n + 1 : q)]; };
    const double v = w0 * cc(i - 1) + w1 * cc(i) + w2 * cc(i + 1) + w3 * cc(i + 2);
    out[g] = (float)(v * scale);
}

// The pivot sequence of the spline system, cp_0 = 1/4, cp_i = 1/(4 -
// cp_{i-1}), i < m: one thread, the same divisions the host makes.
__global__ void k_rs_pivots(double* __restrict__ cp, int m) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double c = 1.0 / 4.0;
    cp[0] = c;
    for (int i = 1; i < m; ++i) {
        c = 1.0 / (4.0 - c);
        cp[i] = c;
    }
}

// Device scratch of adf_resample_f32_noalloc: the pivot sequence, then the
// spline coefficients of every column.  32 MiB (4 Mi doubles) on every
// device the library runs on, held until the process ends.  In host code
// `scratch` names a host-side stand-in, not device memory: the host uses
// only the address hipGetSymbolAddress returns for each device.
constexpr int64_t RS_SCRATCH_DOUBLES = (int64_t)1 << 22;
constexpr int RS_MAX_DEVICES = 64;
static __device__ __align__(256) double scratch[RS_SCRATCH_DOUBLES];
std::mutex g_scratch_mu;
double* g_scratch[RS_MAX_DEVICES] = {};

// Doubles of scratch before the coefficients (the pivots, padded to 256 B).
int64_t rs_pivot_span(int64_t n) { return (std::max<int64_t>(n - 2, 1) + 31) / 32 * 32; }

}  // namespace

// Looks up this device's copy of the resample scratch once; device is the
// current device.  Called by adf_ctx_create.
int adf::resample_scratch_init(int device) {
    if (device < 0 || device >= RS_MAX_DEVICES) return fail(ADF_EINVAL, "resample scratch: device index too large");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[device]) return ADF_OK;
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch[device] = (double*)p;
    return ADF_OK;
}

// The spline resample of ncol columns (see k_rs_coef) on stream st into
// host out (n_out samples per column, column after column).
template <typename T>
static int resample_columns(hipStream_t st, const T* a0, const T* a1, int64_t first, int64_t nsplit, int64_t n,
                            int64_t ncol, int denoise, double t_in0, double dt_in, double t_out0, double dt_out,
                            int64_t n_out, double scale, float* out) {
    std::vector<double> cp((size_t)std::max<int64_t>(n - 2, 1));
    cp[0] = 1.0 / 4.0;
    for (size_t i = 1; i < cp.size(); ++i) cp[i] = 1.0 / (4.0 - cp[i - 1]);
    DevBuf dcp, dcoef, dout;
    HIPCHK(hipMalloc(&dcp.p, cp.size() * 8));
    HIPCHK(hipMalloc(&dcoef.p, (size_t)ncol * (n + 2) * 8));
    HIPCHK(hipMalloc(&dout.p, (size_t)ncol * n_out * 4));
    HIPCHK(hipMemcpyAsync(dcp.p, cp.data(), cp.size() * 8, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL((k_rs_coef<T>), dim3((unsigned)((ncol + 63) / 64)), dim3(64), 0, st, a0, a1, first,
                       (int)nsplit, (int)n, (int)ncol, denoise, (const double*)dcp.p, (double*)dcoef.p);
    const int64_t tot = ncol * n_out;
    hipLaunchKernelGGL(k_rs_eval, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, (const double*)dcoef.p,
                       (int)n, (int)ncol, t_in0, dt_in, t_out0, dt_out, (int)n_out, scale, (float*)dout.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, dout.p, (size_t)tot * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return ADF_OK;
}

extern "C" int adf_solver_gf_resample(adf_solver* s, int64_t first, int64_t n_rec, double timeshift, int64_t nt_gf,
                                      float* out_p, float* out_vn) {
    if (!s) return fail(ADF_EINVAL, "adf_solver_gf_resample: null solver");
    if (first < 0 || n_rec < 1 || first + n_rec > s->ntx)
        return fail(ADF_EINVAL, "adf_solver_gf_resample: transceivers [first, first + n_rec) outside the solver's list");
    if (nt_gf < 1 || !out_p || !out_vn) return fail(ADF_EINVAL, "adf_solver_gf_resample: bad nt_gf or output");
    const int64_t nt = s->dom.nt;
    if (2 * n_rec * (nt + 2) > ((int64_t)1 << 31) || 2 * n_rec * nt_gf > ((int64_t)1 << 31))
        return fail(ADF_EINVAL, "adf_solver_gf_resample: too many samples");
    HIPCHK(hipSetDevice(s->ctx->device));
    hipStream_t st = s->ctx->stream;
    // t_shifted = range(-timeshift, tmax - timeshift; length = nt),
    // t_gf = range(0, tmax; length = nt_gf)
    const double tmax = s->dom.tmax;
    const double dt_in = nt > 1 ? tmax / (double)(nt - 1) : 1.0;
    const double dt_out = nt_gf > 1 ? tmax / (double)(nt_gf - 1) : 1.0;
    const double scale = dt_out / dt_in;
    std::vector<float> out((size_t)2 * n_rec * nt_gf);
    const void* rp = solver_trace_ptr(s, ADF_REC_P);
    const void* rv = solver_trace_ptr(s, ADF_VN_REC);
    int rc;
    if (s->dtype == ADF_F64)
        rc = resample_columns<double>(st, (const double*)rp, (const double*)rv, first, n_rec, nt, 2 * n_rec, 1,
                                      -timeshift, dt_in, 0.0, dt_out, nt_gf, scale, out.data());
    else
        rc = resample_columns<float>(st, (const float*)rp, (const float*)rv, first, n_rec, nt, 2 * n_rec, 1,
                                     -timeshift, dt_in, 0.0, dt_out, nt_gf, scale, out.data());
    if (rc) return rc;
    std::copy(out.begin(), out.begin() + n_rec * nt_gf, out_p);
    std::copy(out.begin() + n_rec * nt_gf, out.end(), out_vn);
    return ADF_OK;
}

extern "C" int adf_resample_f32(adf_ctx* ctx, const float* in, int64_t n_in, int64_t ncol, double t_in0,
                                double dt_in, double t_out0, double dt_out, int64_t n_out, double amp_scale,
                                int denoise, float* out) {
    if (!ctx) return fail(ADF_EINVAL, "adf_resample_f32: null context");
    if (!in || !out || n_in < 1 || ncol < 1 || n_out < 1)
        return fail(ADF_EINVAL, "adf_resample_f32: empty or missing input/output");
    if (!(dt_in > 0.0) || !(dt_out > 0.0)) return fail(ADF_EINVAL, "adf_resample_f32: dt_in and dt_out must be > 0");
    if (ncol * (n_in + 2) > ((int64_t)1 << 31) || ncol * n_out > ((int64_t)1 << 31) || ncol > (1 << 30))
        return fail(ADF_EINVAL, "adf_resample_f32: too many samples");
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    DevBuf din;
    HIPCHK(hipMalloc(&din.p, (size_t)ncol * n_in * 4));
    HIPCHK(hipMemcpyAsync(din.p, in, (size_t)ncol * n_in * 4, hipMemcpyHostToDevice, st));
    return resample_columns<float>(st, (const float*)din.p, (const float*)din.p, 0, ncol, n_in, ncol, denoise ? 1 : 0,
                                   t_in0, dt_in, t_out0, dt_out, n_out, amp_scale, out);
}

extern "C" int adf_resample_f32_noalloc(adf_ctx* ctx, const float* in, int64_t n_in, int64_t ncol, double t_in0,
                                        double dt_in, double t_out0, double dt_out, int64_t n_out, double amp_scale,
                                        int denoise, float* out) {
    if (!ctx) return fail(ADF_EINVAL, "adf_resample_f32_noalloc: null context");
    if (!in || !out || n_in < 1 || ncol < 1 || n_out < 1)
        return fail(ADF_EINVAL, "adf_resample_f32_noalloc: empty or missing input/output");
    if (!(dt_in > 0.0) || !(dt_out > 0.0))
        return fail(ADF_EINVAL, "adf_resample_f32_noalloc: dt_in and dt_out must be > 0");
    if (ncol * n_out > ((int64_t)1 << 31) || ncol > (1 << 30) || n_in > (1 << 30))
        return fail(ADF_EINVAL, "adf_resample_f32_noalloc: too many samples");
    const int64_t span = rs_pivot_span(n_in);
    if (span + ncol * (n_in + 2) > RS_SCRATCH_DOUBLES)
        return fail(ADF_EINVAL, "adf_resample_f32_noalloc: ncol * (n_in + 2) + " + std::to_string(span) +
                                    " doubles exceed the 32 MiB scratch (" + std::to_string(RS_SCRATCH_DOUBLES) +
                                    " doubles)");
    double* base = ctx->device >= 0 && ctx->device < RS_MAX_DEVICES ? g_scratch[ctx->device] : nullptr;
    if (!base) return fail(ADF_EDEVICE, "adf_resample_f32_noalloc: scratch not set up for this device");
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    double* cp = base;
    double* coef = base + span;
    // every pivot and coefficient the evaluation reads is written first:
    // nothing in the scratch needs clearing
    hipLaunchKernelGGL(k_rs_pivots, dim3(1), dim3(1), 0, st, cp, (int)std::max<int64_t>(n_in - 2, 1));
    hipLaunchKernelGGL((k_rs_coef<float>), dim3((unsigned)((ncol + 63) / 64)), dim3(64), 0, st, in, in, (int64_t)0,
                       (int)ncol, (int)n_in, (int)ncol, denoise ? 1 : 0, (const double*)cp, coef);
    const int64_t tot = ncol * n_out;
    hipLaunchKernelGGL(k_rs_eval, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, (const double*)coef,
                       (int)n_in, (int)ncol, t_in0, dt_in, t_out0, dt_out, (int)n_out, amp_scale, out);
    HIPCHK(hipGetLastError());
    return ADF_OK;
}
