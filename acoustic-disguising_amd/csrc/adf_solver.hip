// Single-solve engine: device allocation, transceiver scheduling and the
// launch sequence of one forward_onestep! (src/kernels/forward.jl:13-139).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "adf_internal.hpp"
#include "adf_kernels.hpp"

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
    if ((rc = dmalloc(&s->field, 4 * s->n3 * s->es, "field"))) return rc;
    s->mv_n[0] = p * ny * nz * 2;
    s->mv_n[1] = nx * p * nz * 2;
    s->mv_n[2] = nx * ny * p * 2;
    s->mp_n[0] = q * ny * nz * 2;
    s->mp_n[1] = nx * q * nz * 2;
    s->mp_n[2] = nx * ny * q * 2;
    for (int a = 0; a < 3; ++a) {
        if ((rc = dmalloc(&s->mv[a], s->mv_n[a] * s->es, "cpml.v"))) return rc;
        if ((rc = dmalloc(&s->mp[a], s->mp_n[a] * s->es, "cpml.p"))) return rc;
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
    dfree(s->tx_vn_src); dfree(s->tx_vn_rec); dfree(s->order_p); dfree(s->order_v);
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
    dfree(s->field); dfree(s->field2);
    for (int a = 0; a < 3; ++a) { dfree(s->mv[a]); dfree(s->mp[a]); }
    dfree(s->coef); dfree(s->mask);
    if (s->ev0) hipEventDestroy(s->ev0);
    if (s->ev1) hipEventDestroy(s->ev1);
}

int solver_zero_state(adf_solver* s) {
    hipStream_t st = s->ctx->stream;
    HIPCHK(hipMemsetAsync(s->field, 0, 4 * s->n3 * s->es, st));
    for (int a = 0; a < 3; ++a) {
        if (s->mv_n[a]) HIPCHK(hipMemsetAsync(s->mv[a], 0, s->mv_n[a] * s->es, st));
        if (s->mp_n[a]) HIPCHK(hipMemsetAsync(s->mp[a], 0, s->mp_n[a] * s->es, st));
    }
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

template <typename T>
static Solve<T> make_solve(adf_solver* s) {
    Solve<T> S;
    const int64_t n3 = s->n3, p = s->p, q = p + 1;
    T* F = (T*)s->field;
    S.P = F; S.Vx = F + n3; S.Vy = F + 2 * n3; S.Vz = F + 3 * n3;
    S.mvx = (T*)s->mv[0]; S.mvy = (T*)s->mv[1]; S.mvz = (T*)s->mv[2];
    S.mpx = (T*)s->mp[0]; S.mpy = (T*)s->mp[1]; S.mpz = (T*)s->mp[2];
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
static Txs<T> make_txs(adf_solver* s) {
    Txs<T> X;
    X.idx = s->tx_idx; X.w = (const T*)s->tx_w; X.src = (T*)s->tx_src; X.rec = (T*)s->tx_rec;
    X.dir = (const T*)s->tx_dir; X.vn_src = (T*)s->tx_vn_src; X.vn_rec = (T*)s->tx_vn_rec;
    X.ntx = (int)s->ntx; X.nt = (int)s->dom.nt;
    return X;
}

static constexpr int KZ = 8;

template <typename T, bool EXACT>
static int run_impl(adf_solver* s, int64_t start, int64_t end) {
    hipStream_t st = s->ctx->stream;
    Solve<T> S = make_solve<T>(s);
    Txs<T> X = make_txs<T>(s);
    const int nx = S.nx, ny = S.ny, nz = S.nz;
    const dim3 blk(64, 4, 1);
    const dim3 gp((nx - 1 + 63) / 64, (ny + 3) / 4, (nz + KZ - 1) / KZ);
    const dim3 gv((nx + 63) / 64, (ny + 3) / 4, (nz + KZ - 1) / KZ);
    const int nt = (int)s->dom.nt;
    HIPCHK(hipEventRecord(s->ev0, st));
    for (int64_t it = start; it <= end; ++it) {
        const int it0 = (int)(it - 1);
        hipLaunchKernelGGL((k_update_p<T, EXACT>), gp, blk, 0, st, S, KZ);
        if (s->ntx) {
            for (size_t L = 0; L + 1 < s->lev_p.size(); ++L) {
                const int n = s->lev_p[L + 1] - s->lev_p[L];
                if (n > 0)
                    hipLaunchKernelGGL((k_tx_p<T>), dim3((n + 255) / 256), dim3(256), 0, st, S.P, nx, ny, nz, X,
                                       s->order_p + s->lev_p[L], n, it0);
            }
        }
        hipLaunchKernelGGL((k_update_v<T, EXACT>), gv, blk, 0, st, S, KZ);
        if (s->ntx) {
            const int do_vn2v = s->gf_kind == 0 ? 1 : 0;
            for (size_t L = 0; L + 1 < s->lev_v.size(); ++L) {
                const int n = s->lev_v[L + 1] - s->lev_v[L];
                if (n > 0)
                    hipLaunchKernelGGL((k_tx_v<T>), dim3((n + 255) / 256), dim3(256), 0, st, S.Vx, S.Vy, S.Vz, nx, ny,
                                       nz, X, s->order_v + s->lev_v[L], n, it0, do_vn2v);
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
            const int n = std::max(nx * nz, nx * ny);
            hipLaunchKernelGGL((k_mirror<T>), dim3((n + 255) / 256), dim3(256), 0, st, S);
        }
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(s->ev1, st));
    HIPCHK(hipEventSynchronize(s->ev1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    const int64_t nsteps = end - start + 1;
    s->last_step_ms = nsteps > 0 ? ms / (double)nsteps : 0.0;
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
        hipLaunchKernelGGL((k_slice<double>), dim3((n + 255) / 256), dim3(256), 0, st, (const double*)s->field,
                           (double*)buf, nx, ny, nz, axis, (int)(index - 1));
    else
        hipLaunchKernelGGL((k_slice<float>), dim3((n + 255) / 256), dim3(256), 0, st, (const float*)s->field,
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
