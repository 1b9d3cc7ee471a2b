// Device-memory bandwidth probes (measurement support, no counterpart in
// the reference): the ceiling a streaming kernel reaches on this GPU with
// the access instructions of the field-update sweep, to set the sweep's
// measured HBM bytes/s against, and the kernels whose rocprofv3 counter
// readings (FETCH_SIZE, WRITE_SIZE) calibrate the counters to bytes.
//
//   ADF_BW_COPY_BUF    copy, 128-bit buffer loads and stores (the sweep's
//                      row access: descriptor in SGPRs, lane offset in one
//                      VGPR, row offset in an SGPR)
//   ADF_BW_COPY_GLOBAL copy, 128-bit global loads and stores
//   ADF_BW_READ        read only (128-bit buffer loads, one word written per
//                      workgroup)
//   ADF_BW_WRITE       write only (128-bit buffer stores)
//   ADF_BW_COPY_NT     as ADF_BW_COPY_BUF with non-temporal stores
//   ADF_BW_ZMARCH      the fused sweep's access pattern without its
//                      arithmetic, CPML or mask: each buffer holds four
//                      channels of 256-cell rows (256 rows per plane); a
//                      wave owns 2 rows, marches 4 planes in z and per
//                      plane loads rows -1..1 of three channels and -1..2
//                      of the fourth (the next plane's third channel) one
//                      plane ahead, then stores its 2 rows of all four;
//                      launches alternate between the two buffers
//   ADF_BW_ZMARCH_NT   as ADF_BW_ZMARCH with non-temporal stores
//
// The copy, read and write probes move U 16-byte chunks per thread per
// launch; U chunks are loaded before any is stored.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/adfdtd.h"
#include "adf_fused.hpp"
#include "adf_internal.hpp"

using namespace adf;

namespace {

constexpr int BW_THREADS = 256;
constexpr int BW_U = 1;

template <int POL>
__global__ __launch_bounds__(BW_THREADS) void k_bw_copy_buf(const void* a, void* b, uint32_t bytes) {
    const Rsrc src = rsrc_of(a, bytes), dst = rsrc_of(b, bytes);
    const int base = (blockIdx.x * BW_U * BW_THREADS + threadIdx.x) * 16;
    const int soff = 0;
    u32x4 v[BW_U];
#pragma unroll
    for (int u = 0; u < BW_U; ++u) v[u] = __builtin_amdgcn_raw_buffer_load_b128(src, base + u * BW_THREADS * 16, soff, 0);
#pragma unroll
    for (int u = 0; u < BW_U; ++u) store_b128<POL>(v[u], dst, base + u * BW_THREADS * 16, soff);
}

__global__ __launch_bounds__(BW_THREADS) void k_bw_copy_global(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                               int64_t n16) {
    const int64_t base = (int64_t)blockIdx.x * BW_U * BW_THREADS + threadIdx.x;
    uint4 v[BW_U];
#pragma unroll
    for (int u = 0; u < BW_U; ++u) {
        const int64_t c = base + u * BW_THREADS;
        v[u] = c < n16 ? src[c] : make_uint4(0, 0, 0, 0);
    }
#pragma unroll
    for (int u = 0; u < BW_U; ++u) {
        const int64_t c = base + u * BW_THREADS;
        if (c < n16) dst[c] = v[u];
    }
}

__global__ __launch_bounds__(BW_THREADS) void k_bw_read(const void* a, uint32_t bytes, uint32_t* __restrict__ sink) {
    const Rsrc src = rsrc_of(a, bytes);
    const int base = (blockIdx.x * BW_U * BW_THREADS + threadIdx.x) * 16;
    uint32_t acc = 0;
#pragma unroll
    for (int u = 0; u < BW_U; ++u) {
        const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(src, base + u * BW_THREADS * 16, 0, 0);
        acc ^= v[0] ^ v[1] ^ v[2] ^ v[3];
    }
    // one word per workgroup, written only when the data has a set bit in
    // every lane (zero buffers never write): keeps the loads live
    if (__all(acc != 0) && threadIdx.x == 0) sink[blockIdx.x & 1023] = acc;
}

__global__ __launch_bounds__(BW_THREADS) void k_bw_write(void* b, uint32_t bytes, uint32_t val) {
    const Rsrc dst = rsrc_of(b, bytes);
    const int base = (blockIdx.x * BW_U * BW_THREADS + threadIdx.x) * 16;
    const u32x4 v = {val, val, val, val};
#pragma unroll
    for (int u = 0; u < BW_U; ++u) store_b128<0>(v, dst, base + u * BW_THREADS * 16, 0);
}

// One plane of a z-march tile: rows -1..1 of P, Vx and the next plane's
// Vz, rows -1..2 of Vy (16-byte chunks, 4 cells per lane).
struct ZPlane {
    u32x4 p[3], vx[3], vy[4], vz[3];
};

__device__ __forceinline__ u32x4 zsum(u32x4 a, u32x4 b) { return a + b; }

template <int POL>
__global__ __launch_bounds__(64) void k_bw_zmarch(const void* a, void* b, uint32_t chb, int ny, int nz) {
    constexpr int TY = 2, KB = 4, ROW = 1024;
    const int ntj = ny / TY, nt = ntj * (nz / KB);
    const int t = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    if (t >= nt) return;
    const int j0 = (t % ntj) * TY, k0 = (t / ntj) * KB, k1 = k0 + KB;
    const Rsrc src = rsrc_of(a, 4u * chb), dst = rsrc_of(b, 4u * chb);
    const int vl = (int)(threadIdx.x & 63) * 16;
    auto rowb = [&](int j, int k) {
        j = j < 0 ? 0 : (j >= ny ? ny - 1 : j);
        k = k >= nz ? nz - 1 : k;
        return (int)((uint32_t)(j + ny * k) * (uint32_t)ROW);
    };
    auto ch = [&](int c, int rb) { return (int)((uint32_t)rb + (uint32_t)c * chb); };
    auto load = [&](ZPlane& L, int k) {
#pragma unroll
        for (int r = 0; r <= TY + 1; ++r) {
            const int rb = rowb(j0 + r - 1, k);
            L.vy[r] = __builtin_amdgcn_raw_buffer_load_b128(src, vl, ch(2, rb), 0);
            if (r > TY) continue;
            L.p[r] = __builtin_amdgcn_raw_buffer_load_b128(src, vl, rb, 0);
            L.vx[r] = __builtin_amdgcn_raw_buffer_load_b128(src, vl, ch(1, rb), 0);
            L.vz[r] = __builtin_amdgcn_raw_buffer_load_b128(src, vl, ch(3, rowb(j0 + r - 1, k + 1)), 0);
        }
    };
    u32x4 carry[TY + 1] = {};
    // every loaded word feeds a stored one (integer sums: no arithmetic to
    // speak of), so no load is dead
    auto step = [&](ZPlane& L, int k) {
        u32x4 p1[TY + 1];
#pragma unroll
        for (int r = 0; r <= TY; ++r) p1[r] = zsum(zsum(L.p[r], L.vx[r]), zsum(L.vy[r + 1] - L.vy[r], L.vz[r] - carry[r]));
#pragma unroll
        for (int r = 1; r <= TY; ++r) {
            const int rb = rowb(j0 + r - 1, k);
            store_b128<POL>(p1[r], dst, vl, rb);
            store_b128<POL>(L.vx[r] - p1[r], dst, vl, ch(1, rb));
            store_b128<POL>(L.vy[r] - (p1[r] - p1[r - 1]), dst, vl, ch(2, rb));
            store_b128<POL>(L.vz[r] - p1[r], dst, vl, ch(3, rb));
        }
#pragma unroll
        for (int r = 0; r <= TY; ++r) carry[r] = L.vz[r];
    };
    ZPlane B0, B1;
    load(B0, k0);
    int k = k0;
    for (;;) {
        if (k + 1 < k1) load(B1, k + 1);
        __builtin_amdgcn_sched_barrier(0);
        step(B0, k);
        if (++k >= k1) break;
        if (k + 1 < k1) load(B0, k + 1);
        __builtin_amdgcn_sched_barrier(0);
        step(B1, k);
        if (++k >= k1) break;
    }
}

struct DevMem {
    void* p = nullptr;
    ~DevMem() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

extern "C" int adf_bandwidth_probe(adf_ctx* ctx, int kind, int64_t bytes, int reps, double* ms_per_launch) {
    if (!ctx || !ms_per_launch) return fail(ADF_EINVAL, "adf_bandwidth_probe: null argument");
    const int64_t chunk = (int64_t)BW_U * BW_THREADS * 16;
    if (bytes < chunk || bytes % chunk || bytes > ((int64_t)1 << 30))
        return fail(ADF_EINVAL, "adf_bandwidth_probe: bytes must be a multiple of " + std::to_string(chunk) +
                                    " and at most 1 GiB");
    if (reps < 1 || kind < ADF_BW_COPY_BUF || kind > ADF_BW_ZMARCH_NT)
        return fail(ADF_EINVAL, "adf_bandwidth_probe: bad kind or reps");
    const bool zm = kind == ADF_BW_ZMARCH || kind == ADF_BW_ZMARCH_NT;
    // z-march: four channels of 256 rows x nz planes of 1 KB rows, nz a
    // multiple of the 4-plane tiles
    if (zm && bytes % ((int64_t)4 * 1024 * 256 * 4))
        return fail(ADF_EINVAL, "adf_bandwidth_probe: z-march probes take a multiple of 4 MiB");
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    DevMem a, b, sink;
    HIPCHK(hipMalloc(&a.p, (size_t)bytes));
    HIPCHK(hipMalloc(&b.p, (size_t)bytes));
    HIPCHK(hipMalloc(&sink.p, 4096));
    HIPCHK(hipMemsetAsync(a.p, 0x5a, (size_t)bytes, st));
    HIPCHK(hipMemsetAsync(b.p, 0, (size_t)bytes, st));
    const uint32_t nb = (uint32_t)bytes;
    const unsigned grid = (unsigned)(bytes / chunk);
    const uint32_t zchb = (uint32_t)(bytes / 4);          // bytes per channel
    const int zny = 256, znz = (int)(zchb / 1024 / 256);  // 256 rows of 1 KB per plane
    const unsigned zgrid = (unsigned)((zny / 2) * (znz / 4));
    int flip = 0;
    auto launch = [&]() {
        if (zm) {  // alternate between the buffers: each launch reads what the last one wrote
            const void* from = flip ? b.p : a.p;
            void* to = flip ? a.p : b.p;
            flip ^= 1;
            if (kind == ADF_BW_ZMARCH)
                hipLaunchKernelGGL(k_bw_zmarch<0>, dim3(zgrid), dim3(64), 0, st, from, to, zchb, zny, znz);
            else
                hipLaunchKernelGGL(k_bw_zmarch<ST_NT>, dim3(zgrid), dim3(64), 0, st, from, to, zchb, zny, znz);
            return;
        }
        switch (kind) {
            case ADF_BW_COPY_BUF:
                hipLaunchKernelGGL(k_bw_copy_buf<0>, dim3(grid), dim3(BW_THREADS), 0, st, a.p, b.p, nb);
                break;
            case ADF_BW_COPY_GLOBAL:
                hipLaunchKernelGGL(k_bw_copy_global, dim3(grid), dim3(BW_THREADS), 0, st, (const uint4*)a.p,
                                   (uint4*)b.p, bytes / 16);
                break;
            case ADF_BW_READ:
                hipLaunchKernelGGL(k_bw_read, dim3(grid), dim3(BW_THREADS), 0, st, a.p, nb, (uint32_t*)sink.p);
                break;
            case ADF_BW_WRITE:
                hipLaunchKernelGGL(k_bw_write, dim3(grid), dim3(BW_THREADS), 0, st, b.p, nb, 0x3f800000u);
                break;
            default:
                hipLaunchKernelGGL(k_bw_copy_buf<ST_NT>, dim3(grid), dim3(BW_THREADS), 0, st, a.p, b.p, nb);
        }
    };
    launch();  // warm-up
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, st));
    for (int r = 0; r < reps; ++r) launch();
    HIPCHK(hipEventRecord(e1, st));
    HIPCHK(hipGetLastError());
    HIPCHK(wait_event(e1));
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    *ms_per_launch = (double)ms / reps;
    return ADF_OK;
}
