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
//
// Each thread moves U 16-byte chunks per launch; U chunks are loaded
// before any is stored.
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
    if (reps < 1 || kind < ADF_BW_COPY_BUF || kind > ADF_BW_COPY_NT)
        return fail(ADF_EINVAL, "adf_bandwidth_probe: bad kind or reps");
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
    auto launch = [&]() {
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
