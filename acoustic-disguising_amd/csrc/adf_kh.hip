// Post-hoc Kirchhoff-Helmholtz extrapolation with the analytic kernels
// evaluated on the device (include/adfdtd.h adf_kh_analytic_accumulate):
//
//   out[t, s] += alpha * sum_r sum_{tt=1}^{min(t, ntK)} K(tt; r, s) in[t-tt+1, r]
//
// with K one of the four analytic kernels of kh_kernels.jl:106-157 at the
// times t_grid[tt], optionally minus its tau -> inf pedestal K_dc
// (kh_kernels.jl:174-203) plus alpha K_dc[r, s] cumsum(in)[t, r]
// (direct_kh_accumulate!, kh_kernels.jl:256-296).
//
// The dense K (nt x n_outer x n_inner, 8 GB per kernel at 2000 x 2000 x
// 500) is never formed.  The Gaussian-wavelet kernels are
//   f-terms   nonzero only where |tau| = |t - r/c0| is small,
//   F-term    (v_v only) 0 before the arrival window and 1 after it,
// so for each (receiver s, surface point r) the kernel is evaluated on the
// window |tau| <= 10 sigma (into LDS) and convolved there, and the constant
// parts before and after the window (the F step of v_v, the -K_dc of the
// pedestal form) are sums of the input over index ranges: differences of
// its prefix sums.  Outside the window the dropped f-terms are below
// exp(-50) of their peak; erf (A&S 7.1.26, kh_kernels.jl:19-25) rounds to
// exactly +-1 there.
//
// One workgroup = one receiver s and 256 consecutive output times; it
// walks the surface points r in order, so out[t, s] accumulates the r
// contributions in the reference's order.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "../../include/adfdtd.h"
#include "adf_internal.hpp"

using namespace adf;

namespace {

constexpr int KH_T = 256;    // output times per workgroup
constexpr int KH_W = 1024;   // kernel samples staged in LDS per chunk
constexpr double KH_HALF_WIDTH = 10.0;  // window half-width in sigma

__device__ __forceinline__ double kh_erf(double x) {  // kh_kernels.jl:19-25
    const double s = (x > 0) - (x < 0), ax = fabs(x);
    const double t = 1.0 / (1.0 + 0.3275911 * ax);
    const double y = 1.0 - (((((1.061405429 * t - 1.453152027) * t) + 1.421413741) * t - 0.284496736) * t +
                            0.254829592) * t * exp(-ax * ax);
    return s * y;
}
__device__ __forceinline__ double kh_f(double tau, double sg) {  // kh_kernels.jl:26
    return (1 / (sqrt(2 * M_PI) * sg)) * exp(-(tau * tau) / (2 * (sg * sg)));
}
__device__ __forceinline__ double kh_df(double tau, double sg) { return -(tau / (sg * sg)) * kh_f(tau, sg); }
__device__ __forceinline__ double kh_F(double tau, double sg) { return 0.5 * (1 + kh_erf(tau / (sg * sqrt(2.0)))); }

struct Pair {
    double cr, cs, retard, i1, i2, i3;
};

__device__ __forceinline__ Pair pair_of(const double* op, const double* on, int no, int r, const double* ip,
                                        const double* in_, int ni, int s, double c0) {
    Pair P;
    const double dx = ip[s] - op[r];
    const double dy = ip[ni + s] - op[no + r];
    const double dz = ip[2 * ni + s] - op[2 * no + r];
    const double d = sqrt(dx * dx + dy * dy + dz * dz);
    const double ihx = dx / d, ihy = dy / d, ihz = dz / d;
    P.cr = in_[s] * ihx + in_[ni + s] * ihy + in_[2 * ni + s] * ihz;
    P.cs = on[r] * ihx + on[no + r] * ihy + on[2 * no + r] * ihz;
    P.retard = d / c0;
    P.i1 = 1 / (4 * M_PI * d);
    P.i2 = 1 / (4 * M_PI * (d * d));
    P.i3 = 1 / (4 * M_PI * (d * d * d));
    return P;
}

// The named kernel at tau (kh_kernels.jl:140-155), and its pedestal.
template <int WHICH>
__device__ __forceinline__ double kernel_at(const Pair& P, double tau, double sg, double c0, double rho0) {
    if constexpr (WHICH == 0) return kh_df(tau, sg) * P.i1;
    else if constexpr (WHICH == 1) return P.cs * (kh_f(tau, sg) * P.i2 + kh_df(tau, sg) * P.i1 / c0);
    else if constexpr (WHICH == 2) return (P.cr / rho0) * (kh_f(tau, sg) * P.i2 + kh_df(tau, sg) * P.i1 / c0);
    else
        return (P.cr * P.cs / rho0) *
               (2 * kh_F(tau, sg) * P.i3 + 2 * kh_f(tau, sg) * P.i2 / c0 + kh_df(tau, sg) * P.i1 / (c0 * c0));
}

// number of t_grid entries < v (lower) or <= v (upper)
__device__ __forceinline__ int count_below(const double* g, int n, double v, bool inclusive) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (g[mid] < v || (inclusive && g[mid] == v)) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// cum[r][u] = sum_{t'=1}^{u} in[t', r], u = 0..nt (double)
template <typename T>
__global__ void k_kh_prefix(const T* in, int nt, int no, double* cum) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= no) return;
    double acc = 0.0;
    double* c = cum + (size_t)r * (nt + 1);
    c[0] = 0.0;
    for (int u = 1; u <= nt; ++u) {
        acc += (double)in[(size_t)r * nt + (u - 1)];
        c[u] = acc;
    }
}

// T: input/kernel-sample type (float: K rounded to Float32 as the
// reference's eval_one_kernel(; T = Float32), products in float, sums in
// double); PED: pedestal form.
template <typename T, int WHICH, bool PED>
__global__ __launch_bounds__(KH_T) void k_kh_direct(const double* __restrict__ op, const double* __restrict__ on,
                                                    int no, const double* __restrict__ ip,
                                                    const double* __restrict__ inrm, int ni,
                                                    const double* __restrict__ tg, int ntK, double sg, double c0,
                                                    double rho0, double alpha, const T* __restrict__ in, int nt,
                                                    const double* __restrict__ cum, double* __restrict__ out) {
    __shared__ T Ks[KH_W];
    __shared__ T Is[KH_W + KH_T];
    const int s = blockIdx.y;
    const int T0 = blockIdx.x * KH_T + 1;  // first output time of the tile (1-based)
    const int t = T0 + (int)threadIdx.x;
    const bool live = t <= nt;
    const int m = live ? min(t, ntK) : 0;  // tt runs over 1..m
    double total = 0.0;
    for (int r = 0; r < no; ++r) {
        const Pair P = pair_of(op, on, no, r, ip, inrm, ni, s, c0);
        // window [a, b] (1-based, clipped to 1..ntK)
        const int a = count_below(tg, ntK, P.retard - KH_HALF_WIDTH * sg, false) + 1;
        const int b = count_below(tg, ntK, P.retard + KH_HALF_WIDTH * sg, true);
        const double kc = (P.cr * P.cs / rho0) * (2 * P.i3);  // the F = 1 level of v_v; also K_dc
        double acc = 0.0;
        // ---- window part, in chunks of KH_W kernel samples
        for (int c0i = a; c0i <= b; c0i += KH_W) {
            const int c1 = min(b, c0i + KH_W - 1);
            const int w = c1 - c0i + 1;
            __syncthreads();
            for (int j = threadIdx.x; j < w; j += KH_T) {
                double k = kernel_at<WHICH>(P, tg[c0i + j - 1] - P.retard, sg, c0, rho0);
                if constexpr (PED && WHICH == 3) k = (double)(T)k - (double)(T)kc;
                Ks[j] = (T)k;
            }
            // inputs u = t - tt + 1 for t in the tile and tt in [c0i, c1]:
            // u in [T0 - c1 + 1, T0 + KH_T - 1 - c0i + 1]
            const int u0 = T0 - c1 + 1;
            const int nu = KH_T + w - 1;
            for (int j = threadIdx.x; j < nu; j += KH_T) {
                const int u = u0 + j;
                Is[j] = (u >= 1 && u <= nt) ? in[(size_t)r * nt + (u - 1)] : (T)0;
            }
            __syncthreads();
            if (live) {
                const int hi = min(c1, m);
                for (int tt = c0i; tt <= hi; ++tt) {
                    const T prod = Ks[tt - c0i] * Is[t - tt + 1 - u0];
                    acc += (double)prod;
                }
            }
        }
        if (live && m >= 1) {
            const double* cr_ = cum + (size_t)r * (nt + 1);
            auto csum = [&](int lo, int hi) {  // sum of in[u], u in [lo, hi] (1-based), clipped
                lo = max(lo, 1);
                hi = min(hi, nt);
                return hi >= lo ? cr_[hi] - cr_[lo - 1] : 0.0;
            };
            if constexpr (WHICH == 3 && !PED) {
                // after the window (tt in [b+1, m]) the kernel is the F = 1 level
                if (m > b) acc += (double)(T)kc * csum(t - m + 1, t - max(b, 0));
            }
            if constexpr (WHICH == 3 && PED) {
                // before the window the residual K - K_dc is -K_dc (F = 0)
                const int e = min(a - 1, m);
                if (e >= 1) acc += -(double)(T)kc * csum(t - e + 1, t);
            }
        }
        total += alpha * acc;
    }
    if constexpr (WHICH == 3 && PED) {
        // + alpha K_dc[r, s] cumsum(in)[t, r] (kh_kernels.jl:285-294)
        if (live)
            for (int r = 0; r < no; ++r) {
                const Pair P = pair_of(op, on, no, r, ip, inrm, ni, s, c0);
                const double kc = (double)(T)((P.cr * P.cs / rho0) * (2 * P.i3));
                if (kc == 0) continue;
                total += (alpha * kc) * cum[(size_t)r * (nt + 1) + t];
            }
    }
    if (live) out[(size_t)s * nt + (t - 1)] += total;
}

template <typename T, int WHICH, bool PED>
void launch(dim3 g, hipStream_t st, const double* op, const double* on, int no, const double* ip,
            const double* inn, int ni, const double* tg, int ntK, double sg, double c0, double rho0, double alpha,
            const T* in, int nt, const double* cum, double* out) {
    hipLaunchKernelGGL((k_kh_direct<T, WHICH, PED>), g, dim3(KH_T), 0, st, op, on, no, ip, inn, ni, tg, ntK, sg, c0,
                       rho0, alpha, in, nt, cum, out);
}

template <typename T>
void launch_which(int which, bool ped, dim3 g, hipStream_t st, const double* op, const double* on, int no,
                  const double* ip, const double* inn, int ni, const double* tg, int ntK, double sg, double c0,
                  double rho0, double alpha, const T* in, int nt, const double* cum, double* out) {
#define KH_L(W, P) launch<T, W, P>(g, st, op, on, no, ip, inn, ni, tg, ntK, sg, c0, rho0, alpha, in, nt, cum, out)
    if (which == 0) KH_L(0, false);
    else if (which == 1) KH_L(1, false);
    else if (which == 2) KH_L(2, false);
    else if (ped) KH_L(3, true);
    else KH_L(3, false);
#undef KH_L
}

thread_local double g_kh_ms = 0.0;  // device time of the last call on this thread

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

extern "C" int adf_kh_analytic_accumulate(adf_ctx* ctx, int which, int dtype, int64_t n_outer,
                                          const double* outer_pts, const double* outer_nrm, int64_t n_inner,
                                          const double* inner_pts, const double* inner_nrm, const double* t_grid,
                                          int64_t ntK, double sigma_t, double c0, double rho0, int pedestal,
                                          double alpha, int64_t nt, const double* in, double* out) {
    if (!ctx || !outer_pts || !outer_nrm || !inner_pts || !inner_nrm || !t_grid || !in || !out)
        return fail(ADF_EINVAL, "adf_kh_analytic_accumulate: null argument");
    if (which < 0 || which > 3) return fail(ADF_EINVAL, "adf_kh_analytic_accumulate: which must be 0..3");
    if (dtype != ADF_F64 && dtype != ADF_F32 && dtype != ADF_F32_REF)
        return fail(ADF_EINVAL, "adf_kh_analytic_accumulate: bad dtype");
    if (n_outer < 1 || n_inner < 1 || ntK < 1 || nt < 1 || n_inner > 65535 || nt > (1 << 30) || n_outer > (1 << 30))
        return fail(ADF_EINVAL, "adf_kh_analytic_accumulate: bad sizes");
    if (ntK > nt) return fail(ADF_EINVAL, "adf_kh_analytic_accumulate: K longer than output");
    if (!(sigma_t > 0) || !(c0 > 0) || !(rho0 > 0)) return fail(ADF_EINVAL, "adf_kh_analytic_accumulate: sigma, c0, rho0 > 0");
    for (int64_t i = 1; i < ntK; ++i)
        if (!(t_grid[i] > t_grid[i - 1])) return fail(ADF_EINVAL, "adf_kh_analytic_accumulate: t_grid must increase");
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const bool f64 = dtype == ADF_F64;
    const size_t es = f64 ? 8 : 4;
    DevBuf dop, don, dip, dinn, dtg, din, dcum, dout;
    HIPCHK(hipMalloc(&dop.p, n_outer * 3 * 8));
    HIPCHK(hipMalloc(&don.p, n_outer * 3 * 8));
    HIPCHK(hipMalloc(&dip.p, n_inner * 3 * 8));
    HIPCHK(hipMalloc(&dinn.p, n_inner * 3 * 8));
    HIPCHK(hipMalloc(&dtg.p, ntK * 8));
    HIPCHK(hipMalloc(&din.p, (size_t)n_outer * nt * es));
    HIPCHK(hipMalloc(&dcum.p, (size_t)n_outer * (nt + 1) * 8));
    HIPCHK(hipMalloc(&dout.p, (size_t)n_inner * nt * 8));
    HIPCHK(hipMemcpyAsync(dop.p, outer_pts, n_outer * 3 * 8, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(don.p, outer_nrm, n_outer * 3 * 8, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dip.p, inner_pts, n_inner * 3 * 8, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dinn.p, inner_nrm, n_inner * 3 * 8, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dtg.p, t_grid, ntK * 8, hipMemcpyHostToDevice, st));
    if (f64) {
        HIPCHK(hipMemcpyAsync(din.p, in, (size_t)n_outer * nt * 8, hipMemcpyHostToDevice, st));
    } else {
        std::vector<float> f((size_t)n_outer * nt);
        for (size_t i = 0; i < f.size(); ++i) f[i] = (float)in[i];
        HIPCHK(hipMemcpyAsync(din.p, f.data(), f.size() * 4, hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
    }
    HIPCHK(hipMemcpyAsync(dout.p, out, (size_t)n_inner * nt * 8, hipMemcpyHostToDevice, st));
    const int no = (int)n_outer, ni = (int)n_inner, nti = (int)nt, nk = (int)ntK;
    const dim3 g((unsigned)((nt + KH_T - 1) / KH_T), (unsigned)n_inner);
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, st));
    if (f64) {
        hipLaunchKernelGGL((k_kh_prefix<double>), dim3((no + 255) / 256), dim3(256), 0, st, (const double*)din.p, nti,
                           no, (double*)dcum.p);
        launch_which<double>(which, pedestal != 0, g, st, (const double*)dop.p, (const double*)don.p, no,
                             (const double*)dip.p, (const double*)dinn.p, ni, (const double*)dtg.p, nk, sigma_t, c0,
                             rho0, alpha, (const double*)din.p, nti, (const double*)dcum.p, (double*)dout.p);
    } else {
        hipLaunchKernelGGL((k_kh_prefix<float>), dim3((no + 255) / 256), dim3(256), 0, st, (const float*)din.p, nti,
                           no, (double*)dcum.p);
        launch_which<float>(which, pedestal != 0, g, st, (const double*)dop.p, (const double*)don.p, no,
                            (const double*)dip.p, (const double*)dinn.p, ni, (const double*)dtg.p, nk, sigma_t, c0,
                            rho0, alpha, (const float*)din.p, nti, (const double*)dcum.p, (double*)dout.p);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(e1, st));
    HIPCHK(hipMemcpyAsync(out, dout.p, (size_t)n_inner * nt * 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    g_kh_ms = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ADF_OK;
}

extern "C" int adf_kh_last_ms(double* ms) {
    if (!ms) return fail(ADF_EINVAL, "adf_kh_last_ms: null argument");
    *ms = g_kh_ms;
    return ADF_OK;
}
