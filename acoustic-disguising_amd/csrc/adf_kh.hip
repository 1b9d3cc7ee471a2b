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
// The convolution runs on the VALUs as register-blocked FMAs (k_kh_conv
// below); the surface points are split into chunks whose partial totals
// are summed in chunk order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/adfdtd.h"
#include "adf_internal.hpp"

using namespace adf;

namespace {

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

// cum[r][u] = sum_{t'=1}^{u} in[t', r], u = 0..nt (double).  A workgroup
// takes PX_ROWS surface points; their inputs pass through LDS in chunks of
// PX_CH samples (coalesced loads and stores), and one thread per point adds
// its samples in time order, the order of a serial loop.
constexpr int PX_ROWS = 16, PX_CH = 256;

template <typename T>
__global__ __launch_bounds__(64) void k_kh_prefix(const T* in, int nt, int no, double* cum) {
    __shared__ T tile[PX_ROWS][PX_CH + 1];
    __shared__ double ct[PX_ROWS][PX_CH + 1];
    const int lane = (int)threadIdx.x;
    const int r0 = (int)blockIdx.x * PX_ROWS;
    const int nr = min(PX_ROWS, no - r0);
    double acc = 0.0;
    if (lane < nr) cum[(size_t)(r0 + lane) * (nt + 1)] = 0.0;
    for (int u0 = 0; u0 < nt; u0 += PX_CH) {
        const int nu = min(PX_CH, nt - u0);
        __syncthreads();  // the last chunk's ct is stored
        for (int q = 0; q < nr; ++q)
#pragma unroll
            for (int c = 0; c < PX_CH; c += 64)
                tile[q][c + lane] = c + lane < nu ? in[(size_t)(r0 + q) * nt + u0 + c + lane] : (T)0;
        __syncthreads();
        if (lane < nr) {
#pragma unroll 8
            for (int m = 0; m < PX_CH; ++m) {
                acc += (double)tile[lane][m];
                ct[lane][m] = acc;
            }
        }
        __syncthreads();
        for (int q = 0; q < nr; ++q)
#pragma unroll
            for (int c = 0; c < PX_CH; c += 64)
                if (c + lane < nu) cum[(size_t)(r0 + q) * (nt + 1) + 1 + u0 + c + lane] = ct[q][c + lane];
    }
}

// The convolution, register-blocked.  One wave = one receiver s, 64 R
// consecutive output times and a chunk of surface points r.  Lane l owns
// the R consecutive outputs t = T0 + R l + j; for each r the window [a, b]
// is taken in chunks of taps tt = c1 - kk, kk = 0..16R-1: the kernel samples
// Kw[kk] (wave-uniform, read from LDS as broadcasts) and the inputs
// Is[i] = in[T0 - c1 + 1 + i, r] (zero outside 1..nt) are staged in LDS,
// and output j of lane l takes Kw[kk] * Is[R l + j + kk].  The inputs a
// lane needs for R taps lie in two R-long segments of Is: the segments
// alternate as the taps advance, so each group of R taps loads one new
// segment (R LDS reads for R^2 multiply-adds) and its R kernel samples as
// 16-byte broadcasts.  The 64 lanes' segments start in distinct LDS banks:
// R = 17 for double (odd), R = 16 for float with one unused word after
// every 16 input samples.  Float: two blocks of 1024 outputs cover 2000
// steps (1088-output blocks computed 176 outputs past the end), and the
// p_p/p_v/v_p kernels fit three waves per SIMD; double: two waves.
//
// T = float: kernel samples rounded to Float32 (eval_one_kernel(; T =
// Float32)), inputs Float32, products accumulated with Float32 FMAs over
// one (r, s) window and added to a double total; T = double: all double.
// Zero-padded inputs take care of tt > t (u < 1).  Partial totals of the
// chunks of r are summed in chunk order afterwards (k_kh_reduce), so the
// result does not depend on scheduling.
#ifndef ADF_KH_R32
#define ADF_KH_R32 16  // outputs per lane for float: two blocks of 1024 cover 2000 steps with 2 % to spare
#endif
template <typename T>
struct KC {
    static constexpr int R = sizeof(T) == 4 ? ADF_KH_R32 : 17;
    static constexpr int OUT = 64 * R;
    static constexpr int W = 16 * R;  // taps per LDS chunk (a multiple of R): 256 or 272, one window of 20 sigma / dt
    // kernel samples are stored in groups of R taps at a stride of KS floats
    // (16-byte aligned groups: broadcast reads of four taps at once)
    static constexpr int KS = (R + 3) / 4 * 4;
    // input samples: with an even R, one unused word after every R (lane l
    // reads from word l (R + 1) + m, an odd stride: no LDS bank conflicts)
    static constexpr int PAD = R % 2 == 0 ? 1 : 0;
    static constexpr int IS = (OUT + W) / R * (R + PAD);
    static __device__ __forceinline__ int at(int i) { return PAD ? i + i / R : i; }
};

template <typename T, int R>
__device__ __forceinline__ void kc_seg(const T* Is, int base, T (&x)[R]) {
#pragma unroll
    for (int m = 0; m < R; ++m) x[m] = Is[base + m];
}

// The R taps of group g (kk = gR .. gR + R - 1, stored at Kg = Kw + g KS):
// A = Is[R l + gR ..], B = the next segment.
template <typename T, int R>
__device__ __forceinline__ void kc_block(const T* __restrict__ Kg, const T (&A)[R], const T (&B)[R], T (&acc)[R]) {
    constexpr int KS = (R + 3) / 4 * 4;
    T kv[KS];
    if constexpr (sizeof(T) == 4) {
        typedef float F4 __attribute__((ext_vector_type(4)));
#pragma unroll
        for (int q = 0; q < KS; q += 4) {
            const F4 v = *reinterpret_cast<const F4*>(Kg + q);
            kv[q] = v[0], kv[q + 1] = v[1], kv[q + 2] = v[2], kv[q + 3] = v[3];
        }
    } else {
#pragma unroll
        for (int q = 0; q < R; ++q) kv[q] = Kg[q];
    }
#pragma unroll
    for (int q = 0; q < R; ++q) {
#pragma unroll
        for (int j = 0; j < R; ++j) acc[j] = fma(kv[q], j + q < R ? A[j + q] : B[j + q - R], acc[j]);
    }
}

template <typename T, int WHICH, bool PED>
__global__ __launch_bounds__(64) void k_kh_conv(const double* __restrict__ op, const double* __restrict__ on, int no,
                                                const double* __restrict__ ip, const double* __restrict__ inrm, int ni,
                                                const double* __restrict__ tg, int ntK, double sg, double c0,
                                                double rho0, double alpha, const T* __restrict__ in, int nt,
                                                const double* __restrict__ cum, int rpc, double* __restrict__ dst) {
    constexpr int R = KC<T>::R, OUT = KC<T>::OUT, KW = KC<T>::W, KS = KC<T>::KS;
    __shared__ __attribute__((aligned(16))) T Kw[KW / R * KS];
    __shared__ T Is[KC<T>::IS];
    const int lane = (int)threadIdx.x;
    const int s = blockIdx.y;
    const int T0 = blockIdx.z * OUT + 1;  // first output time of the wave (1-based)
    const int r0 = blockIdx.x * rpc, r1 = min(no, r0 + rpc);
    const int tl = T0 + R * lane;         // the lane's first output
    double tot[R];
#pragma unroll
    for (int j = 0; j < R; ++j) tot[j] = 0.0;
    for (int r = r0; r < r1; ++r) {
        const Pair P = pair_of(op, on, no, r, ip, inrm, ni, s, c0);
        // window [a, b] (1-based, clipped to 1..ntK)
        const int a = count_below(tg, ntK, P.retard - KH_HALF_WIDTH * sg, false) + 1;
        const int b = count_below(tg, ntK, P.retard + KH_HALF_WIDTH * sg, true);
        const double kc = (P.cr * P.cs / rho0) * (2 * P.i3);  // the F = 1 level of v_v; also K_dc
        const T* inr = in + (size_t)r * nt;
        T acc[R];
#pragma unroll
        for (int j = 0; j < R; ++j) acc[j] = (T)0;
        for (int c1 = b; c1 >= a; c1 -= KW) {
            const int cl = max(a, c1 - KW + 1);
            const int w = c1 - cl + 1;
            const int ng = (w + R - 1) / R, wp = ng * R;  // taps rounded to whole groups
            __syncthreads();
            for (int kk = lane; kk < wp; kk += 64) {
                T kv = (T)0;
                if (kk < w) {
                    double k = kernel_at<WHICH>(P, tg[c1 - kk - 1] - P.retard, sg, c0, rho0);
                    if constexpr (PED && WHICH == 3) k = (double)(T)k - (double)(T)kc;
                    kv = (T)k;
                }
                Kw[(kk / R) * KS + kk % R] = kv;
            }
            const int ub = T0 - c1 + 1;  // u of Is[0]
            for (int i = lane; i < OUT + wp; i += 64) {
                const int u = ub + i;
                Is[KC<T>::at(i)] = (u >= 1 && u <= nt) ? inr[u - 1] : (T)0;
            }
            __syncthreads();
            // group g reads segments g and g+1 (input samples base + gR ..
            // base + gR + 2R - 1); the last segment read, ng, ends at
            // OUT + wp - 1.  Segment n of lane l starts at word at(R (l + n)).
            T A[R], B[R];
            const int base = R * lane;
            auto seg = [&](int n, T(&x)[R]) { kc_seg<T, R>(Is, KC<T>::at(base + n * R), x); };
            seg(0, A);
            seg(1, B);
            for (int g = 0; g < ng; g += 2) {
                kc_block<T, R>(Kw + g * KS, A, B, acc);
                if (g + 1 >= ng) break;
                seg(g + 2, A);
                kc_block<T, R>(Kw + (g + 1) * KS, B, A, acc);
                if (g + 2 < ng) seg(g + 3, B);
            }
        }
        const double* cr_ = cum + (size_t)r * (nt + 1);
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const int t = tl + j;
            double v = (double)acc[j];
            if (t <= nt) {
                const int m = min(t, ntK);
                auto csum = [&](int lo, int hi) {  // sum of in[u], u in [lo, hi] (1-based), clipped
                    lo = max(lo, 1);
                    hi = min(hi, nt);
                    return hi >= lo ? cr_[hi] - cr_[lo - 1] : 0.0;
                };
                if constexpr (WHICH == 3 && !PED) {
                    // after the window (tt in [b+1, m]) the kernel is the F = 1 level
                    if (m > b) v += (double)(T)kc * csum(t - m + 1, t - max(b, 0));
                }
                if constexpr (WHICH == 3 && PED) {
                    // before the window the residual K - K_dc is -K_dc (F = 0)
                    const int e = min(a - 1, m);
                    if (e >= 1) v += -(double)(T)kc * csum(t - e + 1, t);
                }
            }
            tot[j] += alpha * v;
            if constexpr (WHICH == 3 && PED) {
                // + alpha K_dc[r, s] cumsum(in)[t, r] (kh_kernels.jl:285-294)
                const double kd = (double)(T)kc;
                if (t <= nt && kd != 0) tot[j] += (alpha * kd) * cr_[t];
            }
        }
    }
    // dst: the chunk's partial totals [chunk][s][t] (or out itself when
    // there is one chunk: accumulate)
    double* d = dst + ((size_t)blockIdx.x * ni + s) * nt;
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const int t = tl + j;
        if (t <= nt) {
            if (gridDim.x == 1) d[t - 1] += tot[j];
            else d[t - 1] = tot[j];
        }
    }
}

// out[s][t] += sum over chunks c (in order) of part[c][s][t]
__global__ void k_kh_reduce(const double* __restrict__ part, int nc, int64_t n, double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = 0.0;
    for (int c = 0; c < nc; ++c) v += part[(size_t)c * n + i];
    out[i] += v;
}

template <typename T, int WHICH, bool PED>
void launch(dim3 g, hipStream_t st, const double* op, const double* on, int no, const double* ip,
            const double* inn, int ni, const double* tg, int ntK, double sg, double c0, double rho0, double alpha,
            const T* in, int nt, const double* cum, int rpc, double* dst) {
    hipLaunchKernelGGL((k_kh_conv<T, WHICH, PED>), g, dim3(64), 0, st, op, on, no, ip, inn, ni, tg, ntK, sg, c0,
                       rho0, alpha, in, nt, cum, rpc, dst);
}

template <typename T>
void launch_which(int which, bool ped, dim3 g, hipStream_t st, const double* op, const double* on, int no,
                  const double* ip, const double* inn, int ni, const double* tg, int ntK, double sg, double c0,
                  double rho0, double alpha, const T* in, int nt, const double* cum, int rpc, double* dst) {
#define KH_L(W, P) launch<T, W, P>(g, st, op, on, no, ip, inn, ni, tg, ntK, sg, c0, rho0, alpha, in, nt, cum, rpc, dst)
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
    // waves: (chunk of surface points, receiver, block of 64 R output
    // times); the points are split into chunks until there are about 4096
    // waves
    const int kc_out = f64 ? KC<double>::OUT : KC<float>::OUT;
    const int nblk = (int)((nt + kc_out - 1) / kc_out);
    const int64_t base_waves = (int64_t)ni * nblk;
    int nc = (int)std::min<int64_t>(no, std::max<int64_t>(1, (4096 + base_waves - 1) / base_waves));
    const int rpc = (no + nc - 1) / nc;
    nc = (no + rpc - 1) / rpc;
    DevBuf dpart;
    if (nc > 1) HIPCHK(hipMalloc(&dpart.p, (size_t)nc * n_inner * nt * 8));
    double* dst = nc > 1 ? (double*)dpart.p : (double*)dout.p;
    const dim3 g((unsigned)nc, (unsigned)n_inner, (unsigned)nblk);
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, st));
    if (f64) {
        hipLaunchKernelGGL((k_kh_prefix<double>), dim3((no + PX_ROWS - 1) / PX_ROWS), dim3(64), 0, st, (const double*)din.p, nti,
                           no, (double*)dcum.p);
        launch_which<double>(which, pedestal != 0, g, st, (const double*)dop.p, (const double*)don.p, no,
                             (const double*)dip.p, (const double*)dinn.p, ni, (const double*)dtg.p, nk, sigma_t, c0,
                             rho0, alpha, (const double*)din.p, nti, (const double*)dcum.p, rpc, dst);
    } else {
        hipLaunchKernelGGL((k_kh_prefix<float>), dim3((no + PX_ROWS - 1) / PX_ROWS), dim3(64), 0, st, (const float*)din.p, nti,
                           no, (double*)dcum.p);
        launch_which<float>(which, pedestal != 0, g, st, (const double*)dop.p, (const double*)don.p, no,
                            (const double*)dip.p, (const double*)dinn.p, ni, (const double*)dtg.p, nk, sigma_t, c0,
                            rho0, alpha, (const float*)din.p, nti, (const double*)dcum.p, rpc, dst);
    }
    if (nc > 1) {
        const int64_t n = (int64_t)n_inner * nt;
        hipLaunchKernelGGL(k_kh_reduce, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, (const double*)dpart.p,
                           nc, n, (double*)dout.p);
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
