// Single-solve engine: device allocation, transceiver scheduling and the
// launch sequence of one forward_onestep! (src/kernels/forward.jl:13-139).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "adf_internal.hpp"
#include "adf_kernels.hpp"
#include "adf_fused.hpp"

namespace adf {

template <typename T>
static void to_type(const double* in, T* out, size_t n) {
    for (size_t i = 0; i < n; ++i) out[i] = (T)in[i];
}

static int dmalloc(void** p, size_t bytes, const char* what) {
    if (bytes == 0) {
        *p = nullptr;
        return ADF_OK;
    }
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(ADF_ENOMEM, std::string("hipMalloc ") + what + ": " + hipGetErrorString(e));
    }
    return ADF_OK;
}

int solver_alloc(adf_solver* s, const adf_cpml* cp) {
    const int64_t nx = s->dom.nx, ny = s->dom.ny, nz = s->dom.nz, p = s->p, q = p + 1;
    s->n3 = nx * ny * nz;
    s->nwx = (nx + 31) / 32;
    int rc;
    for (int b = 0; b < 2; ++b)
        if ((rc = dmalloc(&s->fieldb[b], 4 * s->n3 * s->es, "field"))) return rc;
    s->mv_n[0] = p * ny * nz * 2;
    s->mv_n[1] = nx * p * nz * 2;
    s->mv_n[2] = nx * ny * p * 2;
    s->mp_n[0] = q * ny * nz * 2;
    s->mp_n[1] = nx * q * nz * 2;
    s->mp_n[2] = nx * ny * q * 2;
    // the six CPML arrays of one buffer live back to back (mv_x,mv_y,mv_z,
    // mp_x,mp_y,mp_z), the layout the fused kernel indexes
    {
        int64_t off = 0;
        auto pad = [](int64_t v) { return (v + 63) / 64 * 64; };  // 256-byte aligned sub-arrays
        for (int a = 0; a < 3; ++a) { s->psi_off[a] = off; off = pad(off + s->mv_n[a]); }
        for (int a = 0; a < 3; ++a) { s->psi_off[3 + a] = off; off = pad(off + s->mp_n[a]); }
        s->psi_n = off;
    }
    for (int b = 0; b < 2; ++b) {
        if ((rc = dmalloc(&s->psib[b], s->psi_n * s->es, "cpml memory"))) return rc;
        char* base = (char*)s->psib[b];
        for (int a = 0; a < 3; ++a) {
            s->mvb[b][a] = s->mv_n[a] ? base + s->psi_off[a] * s->es : nullptr;
            s->mpb[b][a] = s->mp_n[a] ? base + s->psi_off[3 + a] * s->es : nullptr;
        }
    }
    const size_t ncoef = 4 * p + 4 * q;
    if ((rc = dmalloc(&s->coef, ncoef * s->es, "cpml coefficients"))) return rc;
    std::vector<double> c(ncoef, 0.0);
    if (p > 0) {
        for (int64_t l = 0; l < p; ++l) {
            c[l] = cp->a_l[l];
            c[p + l] = cp->a_r[l];
            c[2 * p + l] = cp->b_l[l];
            c[3 * p + l] = cp->b_r[l];
        }
        for (int64_t l = 0; l < q; ++l) {
            c[4 * p + l] = cp->a_hl[l];
            c[4 * p + q + l] = cp->a_hr[l];
            c[4 * p + 2 * q + l] = cp->b_hl[l];
            c[4 * p + 3 * q + l] = cp->b_hr[l];
        }
    } else if (cp && cp->a_hl && cp->b_hl) {
        // npml == 0 keeps the 1-entry half-grid tables (unused: guard q > 1)
        c[0] = cp->a_hl[0];
        c[1] = cp->a_hr ? cp->a_hr[0] : 0.0;
        c[2] = cp->b_hl[0];
        c[3] = cp->b_hr ? cp->b_hr[0] : 0.0;
    }
    if (s->es == 8) {
        HIPCHK(hipMemcpy(s->coef, c.data(), ncoef * 8, hipMemcpyHostToDevice));
    } else {
        std::vector<float> cf(ncoef);
        to_type(c.data(), cf.data(), ncoef);
        HIPCHK(hipMemcpy(s->coef, cf.data(), ncoef * 4, hipMemcpyHostToDevice));
    }
    if ((rc = dmalloc((void**)&s->mask, 4 * s->nwx * ny * nz * 4, "mask"))) return rc;
    HIPCHK(hipEventCreate(&s->ev0));
    HIPCHK(hipEventCreate(&s->ev1));
    return solver_zero_state(s);
}

static void dfree(void* p) {
    if (p) hipFree(p);
}

static void free_txs(adf_solver* s) {
    dfree(s->tx_idx); dfree(s->tx_w); dfree(s->tx_src); dfree(s->tx_rec); dfree(s->tx_dir);
    dfree(s->tx_vn_src); dfree(s->tx_vn_rec); dfree(s->order_p); dfree(s->order_v); dfree(s->fix_cells);
    s->fix_cells = nullptr;
    s->n_fix = 0;
    s->tx_idx = nullptr; s->tx_w = s->tx_src = s->tx_rec = s->tx_dir = s->tx_vn_src = s->tx_vn_rec = nullptr;
    s->order_p = s->order_v = nullptr;
    s->ntx = 0;
    s->lev_p.clear();
    s->lev_v.clear();
}

static void free_gf(adf_solver* s) {
    for (int i = 0; i < 4; ++i) { dfree(s->gf_t[i]); s->gf_t[i] = nullptr; }
    dfree(s->gf_irec); dfree(s->gf_isrc); dfree(s->gf_h0); dfree(s->gf_h1);
    s->gf_irec = s->gf_isrc = nullptr;
    s->gf_h0 = s->gf_h1 = nullptr;
    s->gf_kind = 0;
}

void solver_free(adf_solver* s) {
    free_txs(s);
    free_gf(s);
    for (int b = 0; b < 2; ++b) {
        dfree(s->fieldb[b]);
        dfree(s->psib[b]);
    }
    dfree(s->coef); dfree(s->mask); dfree(s->gen_tiles);
    for (hipEvent_t e : s->sweep_ev) hipEventDestroy(e);
    s->sweep_ev.clear();
    if (s->ev0) hipEventDestroy(s->ev0);
    if (s->ev1) hipEventDestroy(s->ev1);
}

int solver_zero_state(adf_solver* s) {
    hipStream_t st = s->ctx->stream;
    for (int b = 0; b < 2; ++b) {
        HIPCHK(hipMemsetAsync(s->fieldb[b], 0, 4 * s->n3 * s->es, st));
        for (int a = 0; a < 3; ++a) {
            if (s->mv_n[a]) HIPCHK(hipMemsetAsync(s->mvb[b][a], 0, s->mv_n[a] * s->es, st));
            if (s->mp_n[a]) HIPCHK(hipMemsetAsync(s->mpb[b][a], 0, s->mp_n[a] * s->es, st));
        }
    }
    s->cur = 0;
    if (s->ntx) {
        const size_t tr = (size_t)s->ntx * s->dom.nt * s->es;
        HIPCHK(hipMemsetAsync(s->tx_rec, 0, 4 * tr, st));
        HIPCHK(hipMemsetAsync(s->tx_vn_rec, 0, tr, st));
    }
    return ADF_OK;
}

// Conflict-free levels: a transceiver whose stencil shares a cell with an
// earlier one goes one level later, so running the levels in order and the
// members of a level in parallel reproduces the reference's sequential
// loop over transceivers (derivatives.jl:30-60, F4 of the survey).
static void make_levels(int64_t ntx, int nx, int ny, int nz, const std::vector<int32_t>& idx,
                        const int* chans, int nch, std::vector<int32_t>& order, std::vector<int>& off) {
    std::vector<std::unordered_map<int64_t, int>> last(nch);
    std::vector<int> lev((size_t)ntx, 1);
    int maxl = 1;
    for (int64_t is = 0; is < ntx; ++is) {
        int L = 1;
        for (int ci = 0; ci < nch; ++ci) {
            const int32_t* id = &idx[(size_t)3 * (is + ntx * chans[ci])];
            const int I = id[0], J = id[1], K = id[2];
            if (!(I >= 1 && I < nx && J >= 1 && J < ny && K >= 1 && K < nz)) continue;
            for (int q = 0; q < 8; ++q) {
                const int64_t cell = (int64_t)(I - 1 + (q & 1)) + (int64_t)nx * ((J - 1 + ((q >> 1) & 1)) + (int64_t)ny * (K - 1 + (q >> 2)));
                auto it = last[ci].find(cell);
                if (it != last[ci].end()) L = std::max(L, it->second + 1);
            }
        }
        lev[is] = L;
        maxl = std::max(maxl, L);
        for (int ci = 0; ci < nch; ++ci) {
            const int32_t* id = &idx[(size_t)3 * (is + ntx * chans[ci])];
            const int I = id[0], J = id[1], K = id[2];
            if (!(I >= 1 && I < nx && J >= 1 && J < ny && K >= 1 && K < nz)) continue;
            for (int q = 0; q < 8; ++q) {
                const int64_t cell = (int64_t)(I - 1 + (q & 1)) + (int64_t)nx * ((J - 1 + ((q >> 1) & 1)) + (int64_t)ny * (K - 1 + (q >> 2)));
                last[ci][cell] = L;
            }
        }
    }
    off.assign(maxl + 1, 0);
    for (int64_t is = 0; is < ntx; ++is) off[lev[is]]++;
    for (int L = 1; L <= maxl; ++L) off[L] += off[L - 1];
    order.assign((size_t)ntx, 0);
    std::vector<int> pos(off.begin(), off.end() - 1);
    for (int64_t is = 0; is < ntx; ++is) order[pos[lev[is] - 1]++] = (int32_t)is;
}

int solver_set_txs(adf_solver* s, int64_t ntx, const int64_t* idx, const void* w, const void* src,
                   const void* dir, const void* vn_src) {
    hipStreamSynchronize(s->ctx->stream);
    free_txs(s);
    if (ntx == 0) {
        s->fuse_mirror = true;
        return ADF_OK;
    }
    const int nx = (int)s->dom.nx, ny = (int)s->dom.ny, nz = (int)s->dom.nz;
    const int64_t nt = s->dom.nt;
    std::vector<int32_t> id32((size_t)12 * ntx);
    for (size_t i = 0; i < id32.size(); ++i) {
        if (idx[i] < -(int64_t(1) << 30) || idx[i] > (int64_t(1) << 30))
            return fail(ADF_EINVAL, "adf_solver_set_transceivers: index out of range");
        id32[i] = (int32_t)idx[i];
    }
    int rc;
    const size_t es = s->es;
    if ((rc = dmalloc((void**)&s->tx_idx, id32.size() * 4, "tx idx"))) return rc;
    if ((rc = dmalloc(&s->tx_w, (size_t)32 * ntx * es, "tx w"))) return rc;
    if ((rc = dmalloc(&s->tx_src, (size_t)4 * ntx * nt * es, "tx src"))) return rc;
    if ((rc = dmalloc(&s->tx_rec, (size_t)4 * ntx * nt * es, "tx rec"))) return rc;
    if ((rc = dmalloc(&s->tx_dir, (size_t)4 * ntx * es, "tx dir"))) return rc;
    if ((rc = dmalloc(&s->tx_vn_src, (size_t)ntx * nt * es, "tx vn_src"))) return rc;
    if ((rc = dmalloc(&s->tx_vn_rec, (size_t)ntx * nt * es, "tx vn_rec"))) return rc;
    if ((rc = dmalloc((void**)&s->order_p, (size_t)ntx * 4, "tx order"))) return rc;
    if ((rc = dmalloc((void**)&s->order_v, (size_t)ntx * 4, "tx order"))) return rc;
    s->ntx = ntx;
    hipStream_t st = s->ctx->stream;
    HIPCHK(hipMemcpyAsync(s->tx_idx, id32.data(), id32.size() * 4, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(s->tx_w, w, (size_t)32 * ntx * es, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(s->tx_src, src, (size_t)4 * ntx * nt * es, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(s->tx_dir, dir, (size_t)4 * ntx * es, hipMemcpyHostToDevice, st));
    if (vn_src) HIPCHK(hipMemcpyAsync(s->tx_vn_src, vn_src, (size_t)ntx * nt * es, hipMemcpyHostToDevice, st));
    else HIPCHK(hipMemsetAsync(s->tx_vn_src, 0, (size_t)ntx * nt * es, st));
    HIPCHK(hipMemsetAsync(s->tx_rec, 0, (size_t)4 * ntx * nt * es, st));
    HIPCHK(hipMemsetAsync(s->tx_vn_rec, 0, (size_t)ntx * nt * es, st));

    std::vector<int32_t> order;
    const int cp[1] = {0};
    make_levels(ntx, nx, ny, nz, id32, cp, 1, order, s->lev_p);
    HIPCHK(hipMemcpyAsync(s->order_p, order.data(), (size_t)ntx * 4, hipMemcpyHostToDevice, st));
    std::vector<int32_t> order_v;
    const int cv[3] = {1, 2, 3};
    make_levels(ntx, nx, ny, nz, id32, cv, 3, order_v, s->lev_v);
    HIPCHK(hipMemcpyAsync(s->order_v, order_v.data(), (size_t)ntx * 4, hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));

    // Velocity cells whose update reads a pressure-injection corner: the
    // fused pass recomputes them after the injection (k_fixup_v).
    {
        std::vector<uint32_t> cells;
        cells.reserve((size_t)ntx * 48);
        for (int64_t is = 0; is < ntx; ++is) {
            const int32_t* id = &id32[(size_t)3 * is];
            const int I = id[0], J = id[1], K = id[2];
            if (!(I >= 1 && I < nx && J >= 1 && J < ny && K >= 1 && K < nz)) continue;
            for (int q = 0; q < 8; ++q) {
                const int i = I - 1 + (q & 1), j = J - 1 + ((q >> 1) & 1), k = K - 1 + (q >> 2);
                const uint32_t c = (uint32_t)(i + nx * (j + ny * k));
                if (i >= 1) cells.push_back(c);
                if (i + 1 <= nx - 1) cells.push_back(c + 1);
                if (j >= 1) cells.push_back((1u << 30) | c);
                if (j + 1 <= ny - 1) cells.push_back((1u << 30) | (c + nx));
                if (k >= 1) cells.push_back((2u << 30) | c);
                if (k + 1 <= nz - 1) cells.push_back((2u << 30) | (c + (uint32_t)(nx * ny)));
            }
        }
        std::sort(cells.begin(), cells.end());
        cells.erase(std::unique(cells.begin(), cells.end()), cells.end());
        s->n_fix = (int)cells.size();
        if (s->n_fix) {
            if ((rc = dmalloc((void**)&s->fix_cells, cells.size() * 4, "fixup list"))) return rc;
            HIPCHK(hipMemcpyAsync(s->fix_cells, cells.data(), cells.size() * 4, hipMemcpyHostToDevice, st));
            HIPCHK(hipStreamSynchronize(st));
        }
    }

    // The mirrors can ride inside the V update unless a velocity stencil
    // touches Vy planes j<=2 or Vz planes k<=2 (1-based).
    s->fuse_mirror = true;
    for (int64_t is = 0; is < ntx; ++is) {
        for (int ch = 2; ch <= 3; ++ch) {
            const int32_t* id = &id32[(size_t)3 * (is + ntx * ch)];
            const int I = id[0], J = id[1], K = id[2];
            if (!(I >= 1 && I < nx && J >= 1 && J < ny && K >= 1 && K < nz)) continue;
            if (ch == 2 && J <= 2) s->fuse_mirror = false;
            if (ch == 3 && K <= 2) s->fuse_mirror = false;
        }
    }
    return ADF_OK;
}

void* solver_trace_ptr(adf_solver* s, int which) {
    if (s->ntx == 0) return nullptr;
    const size_t tr = (size_t)s->ntx * s->dom.nt * s->es;
    if (which >= 0 && which <= 3) return (char*)s->tx_rec + which * tr;
    if (which == 4) return s->tx_vn_rec;
    if (which >= 5 && which <= 8) return (char*)s->tx_src + (which - 5) * tr;
    if (which == 9) return s->tx_vn_src;
    return nullptr;
}

int solver_set_gf(adf_solver* s, int kind, const double* const* tensors, int64_t nt_gf, int64_t dim2,
                  int64_t dim3, const int64_t* iRec, int64_t n_irec, const int64_t* iSrc, int64_t n_isrc,
                  int writeback) {
    hipStreamSynchronize(s->ctx->stream);
    free_gf(s);
    if (kind == 0) return ADF_OK;
    if (kind < 1 || kind > 3 || !tensors || nt_gf < 1 || !iRec || !iSrc)
        return fail(ADF_EINVAL, "adf_solver_set_gf_map: bad kind or null arrays");
    if (kind == 1 && (dim2 != n_isrc || dim3 != n_irec))
        return fail(ADF_EINVAL, "PVFromPin tensors must be (nt_gf, n_isrc, n_irec)");
    if ((kind == 2 || kind == 3) && (dim2 != n_irec || dim3 != n_isrc))
        return fail(ADF_EINVAL, "PVFromPV / PFromPin tensors must be (nt_gf, n_irec, n_isrc)");
    for (int64_t i = 0; i < n_irec; ++i)
        if (iRec[i] < 1 || iRec[i] > s->ntx) return fail(ADF_EINVAL, "iRec out of transceiver range");
    for (int64_t i = 0; i < n_isrc; ++i)
        if (iSrc[i] < 1 || iSrc[i] > s->ntx) return fail(ADF_EINVAL, "iSrc out of transceiver range");
    const int nten = kind == 2 ? 4 : 2;
    const size_t n = (size_t)nt_gf * dim2 * dim3;
    int rc;
    for (int t = 0; t < nten; ++t) {
        if (!tensors[t]) return fail(ADF_EINVAL, "null GF tensor");
        if ((rc = dmalloc(&s->gf_t[t], n * s->es, "GF tensor"))) return rc;
        if (s->es == 8) {
            HIPCHK(hipMemcpy(s->gf_t[t], tensors[t], n * 8, hipMemcpyHostToDevice));
        } else {
            std::vector<float> f(n);
            to_type(tensors[t], f.data(), n);
            HIPCHK(hipMemcpy(s->gf_t[t], f.data(), n * 4, hipMemcpyHostToDevice));
        }
    }
    std::vector<int32_t> ir(n_irec), isr(n_isrc);
    for (int64_t i = 0; i < n_irec; ++i) ir[i] = (int32_t)(iRec[i] - 1);
    for (int64_t i = 0; i < n_isrc; ++i) isr[i] = (int32_t)(iSrc[i] - 1);
    if ((rc = dmalloc((void**)&s->gf_irec, ir.size() * 4, "iRec"))) return rc;
    if ((rc = dmalloc((void**)&s->gf_isrc, isr.size() * 4, "iSrc"))) return rc;
    HIPCHK(hipMemcpy(s->gf_irec, ir.data(), ir.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(s->gf_isrc, isr.data(), isr.size() * 4, hipMemcpyHostToDevice));
    if ((rc = dmalloc((void**)&s->gf_h0, (size_t)n_irec * s->dom.nt * 8, "GF history"))) return rc;
    HIPCHK(hipMemset(s->gf_h0, 0, (size_t)n_irec * s->dom.nt * 8));
    if (kind == 2) {
        if ((rc = dmalloc((void**)&s->gf_h1, (size_t)n_irec * s->dom.nt * 8, "GF history"))) return rc;
        HIPCHK(hipMemset(s->gf_h1, 0, (size_t)n_irec * s->dom.nt * 8));
    }
    s->gf_kind = kind;
    s->gf_wb = writeback ? 1 : 0;
    s->gf_ng = nt_gf;
    s->gf_d2 = dim2;
    s->gf_d3 = dim3;
    s->gf_nr = n_irec;
    s->gf_ns = n_isrc;
    return ADF_OK;
}

void* cur_field(adf_solver* s) { return s->fieldb[s->cur]; }
void* cur_mv(adf_solver* s, int a) { return s->mvb[s->cur][a]; }
void* cur_mp(adf_solver* s, int a) { return s->mpb[s->cur][a]; }

int solver_set_option(adf_solver* s, int option, int64_t value) {
    if (option == ADF_OPT_PATH) {
        if (value < 0 || value > 2) return fail(ADF_EINVAL, "ADF_OPT_PATH must be 0, 1 or 2");
        if (value == 2 && !fused_supported(s)) return fail(ADF_EINVAL, "fused path needs 2 <= nx <= 256");
        s->path = (int)value;
    } else if (option == ADF_OPT_TILE_Y) {
        if (value != 0 && value != 2 && value != 4 && value != 8) return fail(ADF_EINVAL, "ADF_OPT_TILE_Y must be 0, 2, 4 or 8");
        s->tile_y = (int)value;
    } else if (option == ADF_OPT_TIME_SWEEP) {
        s->time_sweep = value != 0;
    } else if (option == ADF_OPT_BLOCK_Z) {
        if (value < 0 || value > (1 << 20)) return fail(ADF_EINVAL, "ADF_OPT_BLOCK_Z out of range");
        s->block_z = (int)value;
    } else {
        return fail(ADF_EINVAL, "unknown option");
    }
    return ADF_OK;
}

template <typename T>
static Solve<T> make_solve(adf_solver* s) {
    Solve<T> S;
    const int64_t n3 = s->n3, p = s->p, q = p + 1;
    T* F = (T*)cur_field(s);
    S.P = F; S.Vx = F + n3; S.Vy = F + 2 * n3; S.Vz = F + 3 * n3;
    S.mvx = (T*)cur_mv(s, 0); S.mvy = (T*)cur_mv(s, 1); S.mvz = (T*)cur_mv(s, 2);
    S.mpx = (T*)cur_mp(s, 0); S.mpy = (T*)cur_mp(s, 1); S.mpz = (T*)cur_mp(s, 2);
    const T* c = (const T*)s->coef;
    S.a = c; S.b = c + 2 * p; S.ah = c + 4 * p; S.bh = c + 4 * p + 2 * q;
    S.mask = s->has_mask ? s->mask : nullptr;
    S.nx = (int)s->dom.nx; S.ny = (int)s->dom.ny; S.nz = (int)s->dom.nz;
    S.p = (int)p; S.nwx = (int)s->nwx;
    S.cpml_p = p > 0 && (s->has_mask || p > 1);
    S.fuse_mirror = s->fuse_mirror ? 1 : 0;
    S.dx = s->dom.dx; S.dy = s->dom.dy; S.dz = s->dom.dz;
    S.fq = s->dom.fact_m0; S.fv = s->dom.fact_m1;
    S.rdx = (float)(1.0 / s->dom.dx); S.rdy = (float)(1.0 / s->dom.dy); S.rdz = (float)(1.0 / s->dom.dz);
    S.fqf = (float)s->dom.fact_m0; S.fvf = (float)s->dom.fact_m1;
    return S;
}

template <typename T>
static Fused<T> make_fused(adf_solver* s, int ty, int kb) {
    Fused<T> F;
    const int nb = s->cur ^ 1;
    F.f0 = (const T*)s->fieldb[s->cur];
    F.f1 = (T*)s->fieldb[nb];
    F.psi0 = (const T*)s->psib[s->cur];
    F.psi1 = (T*)s->psib[nb];
    F.coef = (const T*)s->coef;
    F.mask = s->has_mask ? s->mask : nullptr;
    F.nx = (int)s->dom.nx; F.ny = (int)s->dom.ny; F.nz = (int)s->dom.nz;
    F.p = (int)s->p; F.nwx = (int)s->nwx; F.n3 = (int)s->n3;
    F.cpml_p = s->p > 0 && (s->has_mask || s->p > 1);
    F.fuse_mirror = s->fuse_mirror ? 1 : 0;
    for (int a = 0; a < 3; ++a) {
        F.off_mv[a] = (int)s->psi_off[a];
        F.off_mp[a] = (int)s->psi_off[3 + a];
    }
    F.ty = ty; F.kb = kb;
    F.tiles_y = (F.ny + ty - 1) / ty;
    F.tiles_z = (F.nz + kb - 1) / kb;
    F.dx = s->dom.dx; F.dy = s->dom.dy; F.dz = s->dom.dz;
    F.fq = s->dom.fact_m0; F.fv = s->dom.fact_m1;
    F.rdx = (float)(1.0 / s->dom.dx); F.rdy = (float)(1.0 / s->dom.dy); F.rdz = (float)(1.0 / s->dom.dz);
    F.fqf = (float)s->dom.fact_m0; F.fvf = (float)s->dom.fact_m1;
    return F;
}

template <typename T>
static Txs<T> make_txs(adf_solver* s) {
    Txs<T> X;
    X.idx = s->tx_idx; X.w = (const T*)s->tx_w; X.src = (T*)s->tx_src; X.rec = (T*)s->tx_rec;
    X.dir = (const T*)s->tx_dir; X.vn_src = (T*)s->tx_vn_src; X.vn_rec = (T*)s->tx_vn_rec;
    X.ntx = (int)s->ntx; X.nt = (int)s->dom.nt;
    return X;
}

// Cells per lane of the fused kernel: one wave64 spans the whole x-row.
static int fused_cpl(const adf_solver* s) {
    const int64_t nx = s->dom.nx;
    const int cpl = nx > 128 ? 4 : (nx > 64 ? 2 : 1);
    if (nx > 64 * cpl || nx % cpl != 0) return 0;
    return cpl;
}

bool fused_supported(const adf_solver* s) { return fused_cpl(s) > 0; }

// Tile shape: TY rows per wave and KB planes per wave, chosen so the launch
// has several waves per CU while halo work stays small.
void pick_tiles(const adf_solver* s, int* ty, int* kb) {
    int t = s->tile_y ? s->tile_y : 2;
    int k = s->block_z;
    if (!k) {
        const int64_t ny = s->dom.ny, nz = s->dom.nz;
        const int64_t want = 8LL * s->ctx->num_cu;
        const int64_t tiles_y = (ny + t - 1) / t;
        k = 64;
        while (k > 8 && tiles_y * ((nz + k - 1) / k) < want) k /= 2;
    }
    *ty = t;
    *kb = k;
}

static constexpr int KZ = 8;

// The per-step transceiver / extrapolation / mirror launches shared by both
// paths; S points at the buffers holding the new field.
template <typename T>
static void launch_tx_p(adf_solver* s, const Solve<T>& S, const Txs<T>& X, int it0, hipStream_t st) {
    for (size_t L = 0; L + 1 < s->lev_p.size(); ++L) {
        const int n = s->lev_p[L + 1] - s->lev_p[L];
        if (n > 0)
            hipLaunchKernelGGL((k_tx_p<T>), dim3((n + 255) / 256), dim3(256), 0, st, S.P, S.nx, S.ny, S.nz, X,
                               s->order_p + s->lev_p[L], n, it0);
    }
}

template <typename T>
static void launch_after_v(adf_solver* s, const Solve<T>& S, const Txs<T>& X, int64_t it, hipStream_t st) {
    const int it0 = (int)(it - 1);
    const int nt = (int)s->dom.nt;
    if (s->ntx) {
        const int do_vn2v = s->gf_kind == 0 ? 1 : 0;
        for (size_t L = 0; L + 1 < s->lev_v.size(); ++L) {
            const int n = s->lev_v[L + 1] - s->lev_v[L];
            if (n > 0)
                hipLaunchKernelGGL((k_tx_v<T>), dim3((n + 255) / 256), dim3(256), 0, st, S.Vx, S.Vy, S.Vz, S.nx, S.ny,
                                   S.nz, X, s->order_v + s->lev_v[L], n, it0, do_vn2v);
        }
        if (s->gf_kind) {
            const int nr = (int)s->gf_nr;
            hipLaunchKernelGGL((k_gf_history<T>), dim3((nr + 255) / 256), dim3(256), 0, st, X, s->gf_irec, nr,
                               s->gf_kind, s->dom.z0, s->gf_h0, s->gf_h1, it0);
            if (it < nt) {
                const T* g[4] = {(const T*)s->gf_t[0], (const T*)s->gf_t[1], (const T*)s->gf_t[2],
                                 (const T*)s->gf_t[3]};
                hipLaunchKernelGGL((k_gf_gather<T, T>), dim3((unsigned)s->gf_ns), dim3(256), 0, st, X, s->gf_kind,
                                   s->gf_wb, g[0], g[1], g[2], g[3], (int)s->gf_ng, (int)s->gf_d2, (int)s->gf_d3,
                                   s->gf_isrc, (int)s->gf_ns, nr, s->gf_h0, s->gf_h1, s->dom.z0, it0);
                hipLaunchKernelGGL((k_vn2v<T>), dim3((unsigned)((s->ntx + 255) / 256)), dim3(256), 0, st, X, it0);
            }
        }
    }
    if (!S.fuse_mirror) {
        const int n = std::max(S.nx * S.nz, S.nx * S.ny);
        hipLaunchKernelGGL((k_mirror<T>), dim3((n + 255) / 256), dim3(256), 0, st, S);
    }
}

// Split the (y-tile, z-block) grid into the interior product range, whose
// rows and planes (with their halos) avoid every y/z CPML slab, the mirrored
// planes 1,2 and the rigid planes n, and the list of all other tiles.
static int prepare_tiles(adf_solver* s, int ty, int kb) {
    if (s->tl_ty == ty && s->tl_kb == kb && (s->gen_tiles || s->n_gen == 0) && s->tl_ty) return ADF_OK;
    const int ny = (int)s->dom.ny, nz = (int)s->dom.nz, p = (int)s->p;
    const int tiles_y = (ny + ty - 1) / ty, tiles_z = (nz + kb - 1) / kb;
    const int jlo = std::max(p + 1, 2), jhi = std::min(ny - p - 1, ny - 1);
    const int klo = std::max(p + 1, 2), khi = std::min(nz - p - 1, nz - 1);
    auto yint = [&](int t) { const int j0 = t * ty; return j0 - 1 >= jlo && j0 + ty - 1 < jhi; };
    auto zint = [&](int t) { const int k0 = t * kb; return k0 - 1 >= klo && k0 + kb - 1 < khi; };
    int ylo = -1, yhi = -1, zlo = -1, zhi = -1;
    for (int t = 0; t < tiles_y; ++t) if (yint(t)) { if (ylo < 0) ylo = t; yhi = t + 1; }
    for (int t = 0; t < tiles_z; ++t) if (zint(t)) { if (zlo < 0) zlo = t; zhi = t + 1; }
    if (ylo < 0 || zlo < 0) ylo = yhi = zlo = zhi = 0;
    std::vector<uint32_t> gen;
    for (int tz = 0; tz < tiles_z; ++tz)
        for (int t = 0; t < tiles_y; ++t)
            if (!(t >= ylo && t < yhi && tz >= zlo && tz < zhi)) gen.push_back(((uint32_t)tz << 16) | (uint32_t)t);
    if (tiles_y > 65535 || tiles_z > 65535) return fail(ADF_EINVAL, "grid too large for the tile list");
    dfree(s->gen_tiles);
    s->gen_tiles = nullptr;
    s->n_gen = (int)gen.size();
    if (s->n_gen) {
        int rc = dmalloc((void**)&s->gen_tiles, gen.size() * 4, "tile list");
        if (rc) return rc;
        HIPCHK(hipMemcpy(s->gen_tiles, gen.data(), gen.size() * 4, hipMemcpyHostToDevice));
    }
    s->ty_lo = ylo; s->nty = yhi - ylo; s->tz_lo = zlo; s->ntz = zhi - zlo;
    s->tl_ty = ty; s->tl_kb = kb;
    return ADF_OK;
}

template <typename T, bool EXACT, int CPL, int TY>
static void launch_fused_ty(adf_solver* s, const Fused<T>& F, hipStream_t st) {
    const int n_int = s->nty * s->ntz;
    hipLaunchKernelGGL((k_fused<T, EXACT, CPL, TY>), dim3((unsigned)(s->n_gen + n_int)), dim3(64), 0, st, F,
                       (const uint32_t*)s->gen_tiles, s->n_gen, s->ty_lo, s->nty > 0 ? s->nty : 1, s->tz_lo);
}

template <typename T, bool EXACT, int CPL>
static void launch_fused_cpl(adf_solver* s, const Fused<T>& F, hipStream_t st) {
    if (F.ty == 2) launch_fused_ty<T, EXACT, CPL, 2>(s, F, st);
    else if (F.ty == 8) launch_fused_ty<T, EXACT, CPL, 8>(s, F, st);
    else launch_fused_ty<T, EXACT, CPL, 4>(s, F, st);
}

template <typename T, bool EXACT>
static void launch_fused(adf_solver* s, const Fused<T>& F, int cpl, hipStream_t st) {
    if (cpl == 4) launch_fused_cpl<T, EXACT, 4>(s, F, st);
    else if (cpl == 2) launch_fused_cpl<T, EXACT, 2>(s, F, st);
    else launch_fused_cpl<T, EXACT, 1>(s, F, st);
}

// Start/stop events around one field-update launch (ADF_OPT_TIME_SWEEP).
static void sweep_mark(adf_solver* s, hipStream_t st) {
    if (!s->time_sweep) return;
    if ((int)s->sweep_ev.size() <= s->sweep_used) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return;
        s->sweep_ev.push_back(e);
    }
    (void)hipEventRecord(s->sweep_ev[s->sweep_used++], st);
}

// One time step it (1-based) of forward_onestep! on stream st.
template <typename T, bool EXACT>
static void step_impl(adf_solver* s, int64_t it, hipStream_t st, bool fused, int ty, int kb) {
    const int it0 = (int)(it - 1);
    Txs<T> X = make_txs<T>(s);
    if (fused) {
        Fused<T> F = make_fused<T>(s, ty, kb);
        sweep_mark(s, st);
        launch_fused<T, EXACT>(s, F, fused_cpl(s), st);
        sweep_mark(s, st);
        s->cur ^= 1;  // the new buffers become current
        Solve<T> S = make_solve<T>(s);
        if (s->ntx) {
            launch_tx_p<T>(s, S, X, it0, st);
            if (s->n_fix) {
                // recompute the velocity cells that read injected pressure
                // (F still holds old = 0 and new = 1 buffers)
                hipLaunchKernelGGL((k_fixup_v<T, EXACT>), dim3((s->n_fix + 255) / 256), dim3(256), 0, st, F,
                                   s->fix_cells, s->n_fix);
            }
        }
        launch_after_v<T>(s, S, X, it, st);
    } else {
        Solve<T> S = make_solve<T>(s);
        const dim3 blk(64, 4, 1);
        const int kz = s->block_z ? s->block_z : KZ;
        const dim3 gp((S.nx - 1 + 63) / 64, (S.ny + 3) / 4, (S.nz + kz - 1) / kz);
        const dim3 gv((S.nx + 63) / 64, (S.ny + 3) / 4, (S.nz + kz - 1) / kz);
        sweep_mark(s, st);
        hipLaunchKernelGGL((k_update_p<T, EXACT>), gp, blk, 0, st, S, kz);
        sweep_mark(s, st);
        if (s->ntx) launch_tx_p<T>(s, S, X, it0, st);
        sweep_mark(s, st);
        hipLaunchKernelGGL((k_update_v<T, EXACT>), gv, blk, 0, st, S, kz);
        sweep_mark(s, st);
        launch_after_v<T>(s, S, X, it, st);
    }
}

template <typename T, bool EXACT>
static int run_impl(adf_solver* s, int64_t start, int64_t end) {
    hipStream_t st = s->ctx->stream;
    const bool fused = s->path == 2 || (s->path == 0 && fused_supported(s));
    int ty = 4, kb = 16;
    if (fused) {
        pick_tiles(s, &ty, &kb);
        int rc = prepare_tiles(s, ty, kb);
        if (rc) return rc;
    }
    s->sweep_used = 0;
    s->sweep_avg_ms = 0.0;
    s->sweep_n = 0;
    HIPCHK(hipEventRecord(s->ev0, st));
    for (int64_t it = start; it <= end; ++it) {
        step_impl<T, EXACT>(s, it, st, fused, ty, kb);
        if (s->time_sweep && s->sweep_used >= 4096) {  // bound the event pool: flush
            HIPCHK(hipStreamSynchronize(st));
            for (int e = 0; e + 1 < s->sweep_used; e += 2) {
                float m = 0.f;
                if (hipEventElapsedTime(&m, s->sweep_ev[e], s->sweep_ev[e + 1]) == hipSuccess) {
                    s->sweep_avg_ms += m;
                    s->sweep_n++;
                }
            }
            s->sweep_used = 0;
        }
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(s->ev1, st));
    HIPCHK(hipEventSynchronize(s->ev1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    const int64_t nsteps = end - start + 1;
    s->last_step_ms = nsteps > 0 ? ms / (double)nsteps : 0.0;
    if (s->time_sweep) {
        for (int e = 0; e + 1 < s->sweep_used; e += 2) {
            float m = 0.f;
            if (hipEventElapsedTime(&m, s->sweep_ev[e], s->sweep_ev[e + 1]) == hipSuccess) {
                s->sweep_avg_ms += m;
                s->sweep_n++;
            }
        }
        s->sweep_used = 0;
    }
    return ADF_OK;
}

int solver_sweep_stats(adf_solver* s, double* avg_ms, int64_t* n) {
    *avg_ms = s->sweep_n ? s->sweep_avg_ms / (double)s->sweep_n : 0.0;
    *n = s->sweep_n;
    return ADF_OK;
}

int solver_run(adf_solver* s, int64_t start, int64_t end) {
    if (s->dtype == ADF_F64) return run_impl<double, true>(s, start, end);
    if (s->dtype == ADF_F32_REF) return run_impl<float, true>(s, start, end);
    return run_impl<float, false>(s, start, end);
}

template <typename T>
__global__ void k_slice(const T* P, T* out, int nx, int ny, int nz, int axis, int idx) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (axis == 0) {  // (ny,nz)
        if (t >= ny * nz) return;
        const int j = t % ny, k = t / ny;
        out[t] = P[idx + nx * (j + ny * k)];
    } else if (axis == 1) {  // (nx,nz)
        if (t >= nx * nz) return;
        const int i = t % nx, k = t / nx;
        out[t] = P[i + nx * (idx + ny * k)];
    } else {  // (nx,ny)
        if (t >= nx * ny) return;
        out[t] = P[t + nx * ny * idx];
    }
}

int solver_slice(adf_solver* s, int axis, int64_t index, void* out) {
    const int nx = (int)s->dom.nx, ny = (int)s->dom.ny, nz = (int)s->dom.nz;
    const int64_t n = axis == 0 ? (int64_t)ny * nz : (axis == 1 ? (int64_t)nx * nz : (int64_t)nx * ny);
    const int64_t lim = axis == 0 ? nx : (axis == 1 ? ny : nz);
    if (axis < 0 || axis > 2 || index < 1 || index > lim) return fail(ADF_EINVAL, "slice index out of range");
    void* buf = nullptr;
    int rc;
    if ((rc = dmalloc(&buf, n * s->es, "slice"))) return rc;
    hipStream_t st = s->ctx->stream;
    if (s->es == 8)
        hipLaunchKernelGGL((k_slice<double>), dim3((n + 255) / 256), dim3(256), 0, st, (const double*)cur_field(s),
                           (double*)buf, nx, ny, nz, axis, (int)(index - 1));
    else
        hipLaunchKernelGGL((k_slice<float>), dim3((n + 255) / 256), dim3(256), 0, st, (const float*)cur_field(s),
                           (float*)buf, nx, ny, nz, axis, (int)(index - 1));
    hipError_t e = hipMemcpyAsync(out, buf, n * s->es, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    hipFree(buf);
    if (e != hipSuccess) return fail(ADF_EDEVICE, std::string("slice copy: ") + hipGetErrorString(e));
    return ADF_OK;
}

int solver_build_mask_device(adf_solver* s, int shape, const double* params, const double* xc, const double* yc,
                             const double* zc) {
    const int nx = (int)s->dom.nx, ny = (int)s->dom.ny, nz = (int)s->dom.nz, nwx = (int)s->nwx;
    const size_t pl = (size_t)nwx * ny * nz;
    double* dbuf = nullptr;
    uint32_t* solid = nullptr;
    int rc;
    if ((rc = dmalloc((void**)&dbuf, (size_t)(4 + nx + ny + nz) * 8, "mask coords"))) return rc;
    if ((rc = dmalloc((void**)&solid, pl * 4, "mask solid"))) { hipFree(dbuf); return rc; }
    std::vector<double> h(4 + nx + ny + nz, 0.0);
    const int np = shape == 4 ? 3 : 1;
    for (int i = 0; i < np; ++i) h[i] = params[i];
    std::memcpy(&h[4], xc, nx * 8);
    std::memcpy(&h[4 + nx], yc, ny * 8);
    std::memcpy(&h[4 + nx + ny], zc, nz * 8);
    hipStream_t st = s->ctx->stream;
    hipError_t e = hipMemcpyAsync(dbuf, h.data(), h.size() * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_mask_solid, dim3((unsigned)((pl + 255) / 256)), dim3(256), 0, st, solid, shape, dbuf,
                           dbuf + 4, dbuf + 4 + nx, dbuf + 4 + nx + ny, nx, ny, nz, nwx);
        hipLaunchKernelGGL(k_mask_planes, dim3((unsigned)((pl + 255) / 256)), dim3(256), 0, st, s->mask, solid, nx,
                           ny, nz, nwx);
        e = hipStreamSynchronize(st);
    }
    hipFree(dbuf);
    hipFree(solid);
    if (e != hipSuccess) return fail(ADF_EDEVICE, std::string("mask build: ") + hipGetErrorString(e));
    return ADF_OK;
}

}  // namespace adf
