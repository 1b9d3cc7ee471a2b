// adfdtd C ABI (include/adfdtd.h): contexts, solvers, transfers and the
// per-step launch sequence of forward_onestep! (src/kernels/forward.jl).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/adfdtd.h"
#include "adf_internal.hpp"

namespace adf {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

}  // namespace adf

using namespace adf;

extern "C" int adf_version(void) { return ADF_ABI_VERSION; }
extern "C" const char* adf_last_error(void) { return g_err.c_str(); }

extern "C" int adf_ctx_create(int device, adf_ctx** out) {
    if (!out) return fail(ADF_EINVAL, "adf_ctx_create: null out");
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(ADF_EINVAL, "adf_ctx_create: no such device");
    HIPCHK(hipSetDevice(device));
    adf_ctx* c = new adf_ctx();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) {
        e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
        if (e != hipSuccess) hipStreamDestroy(c->stream);
    }
    if (e != hipSuccess) {
        delete c;
        return fail(ADF_EDEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cu = prop.multiProcessorCount;
    const int rc = resample_scratch_init(device);
    if (rc) {
        hipStreamDestroy(c->side);
        hipStreamDestroy(c->stream);
        delete c;
        return rc;
    }
    *out = c;
    return ADF_OK;
}

extern "C" int adf_ctx_destroy(adf_ctx* c) {
    if (!c) return ADF_OK;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    hipStreamSynchronize(c->side);
    hipStreamDestroy(c->side);
    hipStreamDestroy(c->stream);
    device_pool_release(c->device);
    delete c;
    return ADF_OK;
}

extern "C" int adf_ctx_synchronize(adf_ctx* c) {
    if (!c) return fail(ADF_EINVAL, "null ctx");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    return ADF_OK;
}

// ------------------------------------------------------------- solver

static size_t elem_size(int dtype) { return dtype == ADF_F64 ? 8 : 4; }

extern "C" int adf_solver_create(adf_ctx* ctx, const adf_domain* dom, int dtype, const adf_cpml* cp,
                                 adf_solver** out) {
    if (!ctx || !dom || !out) return fail(ADF_EINVAL, "adf_solver_create: null argument");
    if (dtype != ADF_F32 && dtype != ADF_F64 && dtype != ADF_F32_REF)
        return fail(ADF_EINVAL, "adf_solver_create: dtype must be ADF_F32, ADF_F64 or ADF_F32_REF");
    const int64_t nx = dom->nx, ny = dom->ny, nz = dom->nz;
    if (nx < 3 || ny < 3 || nz < 3) return fail(ADF_EINVAL, "adf_solver_create: grid must be at least 3^3");
    if (nx * ny * nz >= (int64_t(1) << 31) / 4)
        return fail(ADF_EINVAL, "adf_solver_create: grid too large for 32-bit cell indices");
    if (dom->nt < 1) return fail(ADF_EINVAL, "adf_solver_create: nt < 1");
    const int64_t p = cp ? cp->npml : 0;
    if (p < 0 || 2 * p + 1 > std::min(nx, std::min(ny, nz)))
        return fail(ADF_EINVAL, "adf_solver_create: npml out of range for this grid");
    if (p > 0 && (!cp->a_l || !cp->a_r || !cp->a_hl || !cp->a_hr || !cp->b_l || !cp->b_r || !cp->b_hl || !cp->b_hr))
        return fail(ADF_EINVAL, "adf_solver_create: CPML coefficient arrays missing");
    HIPCHK(hipSetDevice(ctx->device));
    adf_solver* s = new adf_solver();
    s->ctx = ctx;
    s->dom = *dom;
    s->dtype = dtype;
    s->es = elem_size(dtype);
    s->p = p;
    int rc = solver_alloc(s, cp);
    if (rc) {
        solver_free(s);
        delete s;
        return rc;
    }
    *out = s;
    return ADF_OK;
}

extern "C" int adf_solver_destroy(adf_solver* s) {
    if (!s) return ADF_OK;
    hipSetDevice(s->ctx->device);
    hipStreamSynchronize(s->ctx->stream);
    if (s->batch) batch_remove(s->batch, s);
    solver_free(s);
    delete s;
    return ADF_OK;
}

extern "C" int adf_solver_set_field(adf_solver* s, const void* field) {
    if (!s || !field) return fail(ADF_EINVAL, "adf_solver_set_field: null argument");
    HIPCHK(hipSetDevice(s->ctx->device));
    HIPCHK(hipMemcpyAsync(cur_field(s), field, 4 * s->n3 * s->es, hipMemcpyHostToDevice, s->ctx->stream));
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
    return ADF_OK;
}

extern "C" int adf_solver_get_field(adf_solver* s, void* field) {
    if (!s || !field) return fail(ADF_EINVAL, "adf_solver_get_field: null argument");
    HIPCHK(hipSetDevice(s->ctx->device));
    HIPCHK(hipMemcpyAsync(field, cur_field(s), 4 * s->n3 * s->es, hipMemcpyDeviceToHost, s->ctx->stream));
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
    return ADF_OK;
}

extern "C" int adf_solver_reset(adf_solver* s) {
    if (!s) return fail(ADF_EINVAL, "adf_solver_reset: null solver");
    HIPCHK(hipSetDevice(s->ctx->device));
    int rc = solver_zero_state(s);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
    return ADF_OK;
}

extern "C" int adf_solver_set_cpml_memory(adf_solver* s, const void* const p[3], const void* const v[3]) {
    if (!s || !p || !v) return fail(ADF_EINVAL, "adf_solver_set_cpml_memory: null argument");
    HIPCHK(hipSetDevice(s->ctx->device));
    for (int a = 0; a < 3; ++a) {
        // both buffers: entries the step never touches must agree in each
        for (int b = 0; b < 2; ++b) {
            if (s->mp_n[a] && p[a]) HIPCHK(hipMemcpyAsync(s->mpb[b][a], p[a], s->mp_n[a] * s->es, hipMemcpyHostToDevice, s->ctx->stream));
            if (s->mv_n[a] && v[a]) HIPCHK(hipMemcpyAsync(s->mvb[b][a], v[a], s->mv_n[a] * s->es, hipMemcpyHostToDevice, s->ctx->stream));
        }
    }
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
    return ADF_OK;
}

extern "C" int adf_solver_get_cpml_memory(adf_solver* s, void* const p[3], void* const v[3]) {
    if (!s || !p || !v) return fail(ADF_EINVAL, "adf_solver_get_cpml_memory: null argument");
    HIPCHK(hipSetDevice(s->ctx->device));
    for (int a = 0; a < 3; ++a) {
        if (s->mp_n[a] && p[a]) HIPCHK(hipMemcpyAsync(p[a], cur_mp(s, a), s->mp_n[a] * s->es, hipMemcpyDeviceToHost, s->ctx->stream));
        if (s->mv_n[a] && v[a]) HIPCHK(hipMemcpyAsync(v[a], cur_mv(s, a), s->mv_n[a] * s->es, hipMemcpyDeviceToHost, s->ctx->stream));
    }
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
    return ADF_OK;
}

extern "C" int adf_solver_set_mask(adf_solver* s, const uint8_t* m) {
    if (!s) return fail(ADF_EINVAL, "adf_solver_set_mask: null solver");
    HIPCHK(hipSetDevice(s->ctx->device));
    touch(s);
    s->open_ok = false;
    if (!m) {
        s->has_mask = false;
        return ADF_OK;
    }
    const int64_t nx = s->dom.nx, ny = s->dom.ny, nz = s->dom.nz;
    const int64_t nwx = s->nwx;
    std::vector<uint32_t> bits((size_t)(4 * nwx * ny * nz), 0u);
    for (int c = 0; c < 4; ++c)
        for (int64_t k = 0; k < nz; ++k)
            for (int64_t j = 0; j < ny; ++j) {
                const uint8_t* row = m + c * s->n3 + nx * (j + ny * k);
                uint32_t* w = &bits[(size_t)(((c * nz + k) * ny + j) * nwx)];
                for (int64_t i = 0; i < nx; ++i)
                    if (row[i]) w[i >> 5] |= 1u << (i & 31);
            }
    HIPCHK(hipMemcpyAsync(s->mask, bits.data(), bits.size() * 4, hipMemcpyHostToDevice, s->ctx->stream));
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
    s->has_mask = true;
    return ADF_OK;
}

extern "C" int adf_solver_build_mask(adf_solver* s, int shape, const double* params, const double* xc,
                                     const double* yc, const double* zc) {
    if (!s) return fail(ADF_EINVAL, "adf_solver_build_mask: null solver");
    touch(s);
    s->open_ok = false;
    if (shape == 0) {
        s->has_mask = false;
        return ADF_OK;
    }
    if (shape < 1 || shape > 4 || !params || !xc || !yc || !zc)
        return fail(ADF_EINVAL, "adf_solver_build_mask: bad shape or null coordinates");
    HIPCHK(hipSetDevice(s->ctx->device));
    int rc = solver_build_mask_device(s, shape, params, xc, yc, zc);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
    s->has_mask = true;
    return ADF_OK;
}

extern "C" int adf_solver_get_mask(adf_solver* s, uint8_t* m) {
    if (!s || !m) return fail(ADF_EINVAL, "adf_solver_get_mask: null argument");
    if (!s->has_mask) return fail(ADF_ESTATE, "adf_solver_get_mask: solver has no mask");
    HIPCHK(hipSetDevice(s->ctx->device));
    const int64_t nx = s->dom.nx, ny = s->dom.ny, nz = s->dom.nz, nwx = s->nwx;
    std::vector<uint32_t> bits((size_t)(4 * nwx * ny * nz));
    HIPCHK(hipMemcpyAsync(bits.data(), s->mask, bits.size() * 4, hipMemcpyDeviceToHost, s->ctx->stream));
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
    for (int c = 0; c < 4; ++c)
        for (int64_t k = 0; k < nz; ++k)
            for (int64_t j = 0; j < ny; ++j) {
                uint8_t* row = m + c * s->n3 + nx * (j + ny * k);
                const uint32_t* w = &bits[(size_t)(((c * nz + k) * ny + j) * nwx)];
                for (int64_t i = 0; i < nx; ++i) row[i] = (w[i >> 5] >> (i & 31)) & 1u;
            }
    return ADF_OK;
}

extern "C" int adf_solver_set_transceivers(adf_solver* s, int64_t ntx, const int64_t* idx, const void* w,
                                           const void* src, const void* dir, const void* vn_src) {
    if (!s) return fail(ADF_EINVAL, "adf_solver_set_transceivers: null solver");
    if (ntx < 0 || (ntx > 0 && (!idx || !w || !src || !dir)))
        return fail(ADF_EINVAL, "adf_solver_set_transceivers: null arrays");
    if (ntx > (int64_t(1) << 24)) return fail(ADF_EINVAL, "adf_solver_set_transceivers: too many transceivers");
    HIPCHK(hipSetDevice(s->ctx->device));
    touch(s);
    return solver_set_txs(s, ntx, idx, w, src, dir, vn_src);
}

extern "C" int adf_solver_get_trace(adf_solver* s, int which, void* out) {
    if (!s || !out) return fail(ADF_EINVAL, "adf_solver_get_trace: null argument");
    void* p = solver_trace_ptr(s, which);
    if (!p) return fail(ADF_EINVAL, "adf_solver_get_trace: bad buffer id or no transceivers");
    HIPCHK(hipSetDevice(s->ctx->device));
    HIPCHK(hipMemcpyAsync(out, p, (size_t)s->ntx * s->dom.nt * s->es, hipMemcpyDeviceToHost, s->ctx->stream));
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
    return ADF_OK;
}

extern "C" int adf_solver_set_trace(adf_solver* s, int which, const void* in) {
    if (!s || !in) return fail(ADF_EINVAL, "adf_solver_set_trace: null argument");
    void* p = solver_trace_ptr(s, which);
    if (!p) return fail(ADF_EINVAL, "adf_solver_set_trace: bad buffer id or no transceivers");
    HIPCHK(hipSetDevice(s->ctx->device));
    HIPCHK(hipMemcpyAsync(p, in, (size_t)s->ntx * s->dom.nt * s->es, hipMemcpyHostToDevice, s->ctx->stream));
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
    touch(s);
    return solver_trace_written(s, which, in);
}

extern "C" int adf_solver_set_gf_map(adf_solver* s, int kind, const double* const* tensors, int64_t nt_gf,
                                     int64_t dim2, int64_t dim3, const int64_t* iRec, int64_t n_irec,
                                     const int64_t* iSrc, int64_t n_isrc, int writeback) {
    if (!s) return fail(ADF_EINVAL, "adf_solver_set_gf_map: null solver");
    HIPCHK(hipSetDevice(s->ctx->device));
    touch(s);
    return solver_set_gf(s, kind, tensors, nt_gf, dim2, dim3, iRec, n_irec, iSrc, n_isrc, writeback);
}

extern "C" int adf_solver_run(adf_solver* s, int64_t start_step, int64_t end_step) {
    if (!s) return fail(ADF_EINVAL, "adf_solver_run: null solver");
    if (start_step < 1 || end_step > s->dom.nt || start_step > end_step + 1)
        return fail(ADF_EINVAL, "adf_solver_run: steps must satisfy 1 <= start <= end+1, end <= nt");
    HIPCHK(hipSetDevice(s->ctx->device));
    return solver_run(s, start_step, end_step);
}

extern "C" int adf_solver_get_pressure_slice(adf_solver* s, int axis, int64_t index, void* out) {
    if (!s || !out) return fail(ADF_EINVAL, "adf_solver_get_pressure_slice: null argument");
    HIPCHK(hipSetDevice(s->ctx->device));
    return solver_slice(s, axis, index, out);
}

extern "C" int adf_solver_set_option(adf_solver* s, int option, int64_t value) {
    if (!s) return fail(ADF_EINVAL, "adf_solver_set_option: null solver");
    // options that change the tiling invalidate the tile lists themselves;
    // timing, store hint and transceiver launch form leave them valid
    return solver_set_option(s, option, value);
}

extern "C" int adf_solver_sweep_stats(adf_solver* s, double* avg_ms, int64_t* launches) {
    if (!s || !avg_ms || !launches) return fail(ADF_EINVAL, "adf_solver_sweep_stats: null argument");
    return solver_sweep_stats(s, avg_ms, launches);
}

extern "C" int adf_solver_last_step_ms(adf_solver* s, double* ms) {
    if (!s || !ms) return fail(ADF_EINVAL, "adf_solver_last_step_ms: null argument");
    *ms = s->last_step_ms;
    return ADF_OK;
}

extern "C" int adf_solver_next_step(adf_solver* s, int64_t* it) {
    if (!s || !it) return fail(ADF_EINVAL, "adf_solver_next_step: null argument");
    *it = s->next_it;
    return ADF_OK;
}

// -------------------------------------------------------------- batch

extern "C" int adf_batch_create(adf_ctx* ctx, adf_batch** out) {
    if (!ctx || !out) return fail(ADF_EINVAL, "adf_batch_create: null argument");
    adf_batch* b = new adf_batch();
    b->ctx = ctx;
    *out = b;
    return ADF_OK;
}

extern "C" int adf_batch_destroy(adf_batch* b) {
    if (!b) return ADF_OK;
    hipSetDevice(b->ctx->device);
    hipStreamSynchronize(b->ctx->stream);
    batch_free(b);
    delete b;
    return ADF_OK;
}

extern "C" int adf_batch_add(adf_batch* b, adf_solver* s) {
    if (!b || !s) return fail(ADF_EINVAL, "adf_batch_add: null argument");
    return batch_add(b, s);
}

extern "C" int adf_batch_remove(adf_batch* b, adf_solver* s) {
    if (!b || !s) return fail(ADF_EINVAL, "adf_batch_remove: null argument");
    HIPCHK(hipSetDevice(b->ctx->device));
    return batch_remove(b, s);
}

extern "C" int adf_batch_size(adf_batch* b, int64_t* n) {
    if (!b || !n) return fail(ADF_EINVAL, "adf_batch_size: null argument");
    *n = (int64_t)b->members.size();
    return ADF_OK;
}

extern "C" int adf_batch_run(adf_batch* b, int64_t nsteps) {
    if (!b) return fail(ADF_EINVAL, "adf_batch_run: null batch");
    HIPCHK(hipSetDevice(b->ctx->device));
    return batch_run(b, nsteps);
}

extern "C" int adf_batch_set_option(adf_batch* b, int option, int64_t value) {
    if (!b) return fail(ADF_EINVAL, "adf_batch_set_option: null batch");
    if (option == ADF_BATCH_PERSISTENT) {
        if (value != 0 && value != 1) return fail(ADF_EINVAL, "ADF_BATCH_PERSISTENT: 0 or 1");
        b->persistent = value != 0;
    } else if (option == ADF_BATCH_GRAPH_STEPS) {
        if (value < 0 || value > 4096) return fail(ADF_EINVAL, "ADF_BATCH_GRAPH_STEPS: 0..4096");
        b->graph_steps = (int)value;
    } else {
        return fail(ADF_EINVAL, "adf_batch_set_option: unknown option");
    }
    b->dirty = true;
    return ADF_OK;
}

extern "C" int adf_batch_last_step_ms(adf_batch* b, double* ms) {
    if (!b || !ms) return fail(ADF_EINVAL, "adf_batch_last_step_ms: null argument");
    *ms = b->last_step_ms;
    return ADF_OK;
}
