// Single-solve engine: device allocation, transceiver scheduling and the
// launch sequence of one forward_onestep! (src/kernels/forward.jl:13-139).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "adf_internal.hpp"
#include "adf_kernels.hpp"
#include "adf_fused.hpp"
#include "adf_sweep.hpp"
#include "adf_batch.hpp"
#include "adf_launch.hpp"

namespace adf {

template <typename T>
static void to_type(const double* in, T* out, size_t n) {
    for (size_t i = 0; i < n; ++i) out[i] = (T)in[i];
}

// Device blocks that solvers, batches and servers free are kept per device
// and size for the next allocation of the same size.  hipFree waits for the
// device and costs ~0.1 ms per block, which made creating and destroying
// many solves of one shape (one per source in impulsive_gfs, one per served
// request) a visible part of their time.  hipFree also waits for the device
// to go idle; the pool keeps that guarantee at reuse instead: each pooled
// block carries the count of frees at its free, and a block freed after the
// last device synchronization is handed out only after another one.  The
// pool holds at most POOL_MAX bytes; an allocation that fails releases the
// device's pool and tries again.  Reused blocks are not zeroed: every
// buffer is written or cleared after allocation.
namespace {
std::mutex pool_mu;
std::unordered_map<void*, std::pair<int, size_t>> pool_live;  // block -> (device, bytes)
std::map<std::pair<int, size_t>, std::vector<std::pair<void*, uint64_t>>> pool_free;  // (block, free count)
size_t pool_bytes = 0;                        // all devices (POOL_MAX applies to the sum)
std::unordered_map<int, size_t> pool_dev_bytes;  // per device
uint64_t pool_frees = 0;                      // blocks pooled so far
// per device: pool_frees at the start of its last synchronization
std::unordered_map<int, uint64_t> pool_synced;
constexpr size_t POOL_MAX = (size_t)32 << 30;

// Frees pooled blocks of device dev until at most keep bytes of it stay
// pooled (largest blocks first).
void pool_trim_locked(int dev, size_t keep) {
    for (auto it = pool_free.end(); it != pool_free.begin() && pool_dev_bytes[dev] > keep;) {
        --it;
        if (it->first.first != dev) continue;
        auto& v = it->second;
        while (!v.empty() && pool_dev_bytes[dev] > keep) {
            (void)hipFree(v.back().first);
            v.pop_back();
            pool_bytes -= it->first.second;
            pool_dev_bytes[dev] -= it->first.second;
        }
        if (v.empty()) it = pool_free.erase(it);
    }
}
}  // namespace

// Frees the pooled blocks of device dev.
void device_pool_release(int dev) {
    std::lock_guard<std::mutex> lk(pool_mu);
    pool_trim_locked(dev, 0);
}

// Frees pooled blocks of device dev until at most keep bytes stay pooled.
void device_pool_trim(int dev, size_t keep) {
    std::lock_guard<std::mutex> lk(pool_mu);
    pool_trim_locked(dev, keep);
}

constexpr size_t FIELD_SLACK = 256;
// MI355X's last-level (Infinity) cache
constexpr int64_t LLC_BYTES = int64_t(256) << 20;

static int dmalloc(void** p, size_t bytes, const char* what) {
    *p = nullptr;
    if (bytes == 0) return ADF_OK;
    int dev = 0;
    (void)hipGetDevice(&dev);
    {
        std::unique_lock<std::mutex> lk(pool_mu);
        auto it = pool_free.find({dev, bytes});
        if (it != pool_free.end() && !it->second.empty()) {
            const std::pair<void*, uint64_t> b = it->second.back();
            it->second.pop_back();
            pool_bytes -= bytes;
            pool_dev_bytes[dev] -= bytes;
            pool_live[b.first] = {dev, bytes};
            if (b.second > pool_synced[dev]) {  // freed since the device's last synchronization
                const uint64_t upto = pool_frees;
                lk.unlock();
                (void)hipDeviceSynchronize();
                lk.lock();
                pool_synced[dev] = std::max(pool_synced[dev], upto);
            }
            *p = b.first;
            return ADF_OK;
        }
    }
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        device_pool_release(dev);
        e = hipMalloc(p, bytes);
    }
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(ADF_ENOMEM, std::string("hipMalloc ") + what + ": " + hipGetErrorString(e));
    }
    std::lock_guard<std::mutex> lk(pool_mu);
    pool_live[*p] = {dev, bytes};
    return ADF_OK;
}

int solver_alloc(adf_solver* s, const adf_cpml* cp) {
    const int64_t nx = s->dom.nx, ny = s->dom.ny, nz = s->dom.nz, p = s->p, q = p + 1;
    s->n3 = nx * ny * nz;
    s->nwx = (nx + 31) / 32;
    int rc;
    // FIELD_SLACK bytes past each buffer: the fused sweep's last lane of a
    // row whose length is not a multiple of its cells per lane loads (never
    // stores) up to 3 cells past the row, unchecked at the buffer's end
    // because the row offset is the scalar one
    for (int b = 0; b < 2; ++b)
        if ((rc = dmalloc(&s->fieldb[b], 4 * s->n3 * s->es + FIELD_SLACK, "field"))) return rc;
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
        s->psi_n = off + 64;  // slack after the last array (as FIELD_SLACK)
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
    if (!p) return;
    {
        std::lock_guard<std::mutex> lk(pool_mu);
        auto it = pool_live.find(p);
        if (it != pool_live.end()) {
            const std::pair<int, size_t> key = it->second;
            pool_live.erase(it);
            if (pool_bytes + key.second <= POOL_MAX) {
                pool_free[key].push_back({p, ++pool_frees});
                pool_bytes += key.second;
                pool_dev_bytes[key.first] += key.second;
                return;
            }
        }
    }
    (void)hipFree(p);
}

static void free_txs(adf_solver* s) {
    dfree(s->tx_idx); dfree(s->tx_w); dfree(s->tx_src); dfree(s->tx_rec); dfree(s->tx_dir);
    dfree(s->tx_vn_src); dfree(s->tx_vn_rec); dfree(s->order_p); dfree(s->order_v); dfree(s->fix_cells);
    dfree(s->gf_slot_first); dfree(s->gf_slot_next); dfree(s->txc_list); dfree(s->txc_hdr);
    s->gf_slot_first = s->gf_slot_next = nullptr;
    s->txc_hdr = nullptr;
    s->fix_cells = nullptr;
    s->n_fix = 0;
    s->h_fix.clear();
    s->h_order_p.clear();
    s->h_order_v.clear();
    s->txc_list = nullptr;
    s->txc_ok = false;
    s->tx_idx = nullptr; s->tx_w = s->tx_src = s->tx_rec = s->tx_dir = s->tx_vn_src = s->tx_vn_rec = nullptr;
    s->order_p = s->order_v = nullptr;
    s->ntx = 0;
    s->lev_p.clear();
    s->lev_v.clear();
}

static void free_gf(adf_solver* s) {
    for (int i = 0; i < 4; ++i) {
        dfree(s->gf_t[i]);
        dfree(s->gf_tn[i]);
        s->gf_t[i] = s->gf_tn[i] = nullptr;
    }
    dfree(s->gf_irec); dfree(s->gf_isrc); dfree(s->gf_h0); dfree(s->gf_h1); dfree(s->gf_part);
    s->gf_irec = s->gf_isrc = nullptr;
    s->gf_h0 = s->gf_h1 = s->gf_part = nullptr;
    s->gf_kind = 0;
    s->gf_B = -1;
    s->gf_hist_valid = 0;
    s->gf_max_index = 0;
    s->gf_isrc_h.clear();
    s->gf_irec_h.clear();
}

void solver_free(adf_solver* s) {
    free_txs(s);
    free_gf(s);
    for (int b = 0; b < 2; ++b) {
        dfree(s->fieldb[b]);
        dfree(s->psib[b]);
    }
    dfree(s->coef); dfree(s->mask); for (auto& t : s->tiles) dfree(t);
    dfree(s->txs_tiles); dfree(s->txs_wait); dfree(s->txs_wait_off); dfree(s->txs_flag); dfree(s->txs_err);
    if (s->txs_err_h) (void)hipHostFree(s->txs_err_h);
    if (s->ev_fork) hipEventDestroy(s->ev_fork);
    if (s->ev_join) hipEventDestroy(s->ev_join);
    for (hipEvent_t e : s->sweep_ev) hipEventDestroy(e);
    s->sweep_ev.clear();
    if (s->ev0) hipEventDestroy(s->ev0);
    if (s->ev1) hipEventDestroy(s->ev1);
}

int solver_zero_state(adf_solver* s) {
    hipStream_t st = s->ctx->stream;
    for (int b = 0; b < 2; ++b) {
        HIPCHK(hipMemsetAsync(s->fieldb[b], 0, 4 * s->n3 * s->es, st));
        // the whole CPML buffer, padding included (the interior sweep loads
        // whole rows next to the x slabs)
        if (s->psi_n) HIPCHK(hipMemsetAsync(s->psib[b], 0, s->psi_n * s->es, st));
    }
    s->cur = 0;
    s->next_it = 1;
    if (s->ntx) {
        const size_t tr = (size_t)s->ntx * s->dom.nt * s->es;
        HIPCHK(hipMemsetAsync(s->tx_rec, 0, 4 * tr, st));
        HIPCHK(hipMemsetAsync(s->tx_vn_rec, 0, tr, st));
    }
    return ADF_OK;
}

// Conflict-free levels.  Only transceivers that inject (a nonzero source
// sample on the channel, or sources written by the GF map) change the field;
// the others only record.  A transceiver must run after every earlier
// injector that shares a stencil cell with it, and an injector also after
// every earlier recorder of a shared cell.  Running the levels in order and
// the members of a level in parallel then reproduces the reference's
// sequential loop over transceivers (derivatives.jl:30-60) value for value.
// order entries: transceiver index | (injector << 31).
static void make_levels(int64_t ntx, int nx, int ny, int nz, const std::vector<int32_t>& idx,
                        const std::vector<uint8_t>& inj, const int* chans, int nch, std::vector<uint32_t>& order,
                        std::vector<int>& off) {
    std::vector<std::unordered_map<int64_t, int>> last_any(nch), last_inj(nch);
    std::vector<int> lev((size_t)ntx, 1);
    int maxl = 1;
    auto cells = [&](int64_t is, int ci, int64_t* out) {
        const int32_t* id = &idx[(size_t)3 * (is + ntx * chans[ci])];
        const int I = id[0], J = id[1], K = id[2];
        if (!(I >= 1 && I < nx && J >= 1 && J < ny && K >= 1 && K < nz)) return false;
        for (int q = 0; q < 8; ++q)
            out[q] = (int64_t)(I - 1 + (q & 1)) + (int64_t)nx * ((J - 1 + ((q >> 1) & 1)) + (int64_t)ny * (K - 1 + (q >> 2)));
        return true;
    };
    int64_t cs[8];
    for (int64_t is = 0; is < ntx; ++is) {
        int L = 1;
        const bool I = inj[is] != 0;
        for (int ci = 0; ci < nch; ++ci) {
            if (!cells(is, ci, cs)) continue;
            auto& m = I ? last_any[ci] : last_inj[ci];
            for (int q = 0; q < 8; ++q) {
                auto it = m.find(cs[q]);
                if (it != m.end()) L = std::max(L, it->second + 1);
            }
        }
        lev[is] = L;
        maxl = std::max(maxl, L);
        for (int ci = 0; ci < nch; ++ci) {
            if (!cells(is, ci, cs)) continue;
            for (int q = 0; q < 8; ++q) {
                int& a = last_any[ci][cs[q]];
                a = std::max(a, L);
                if (I) {
                    int& b = last_inj[ci][cs[q]];
                    b = std::max(b, L);
                }
            }
        }
    }
    off.assign(maxl + 1, 0);
    for (int64_t is = 0; is < ntx; ++is) off[lev[is]]++;
    for (int L = 1; L <= maxl; ++L) off[L] += off[L - 1];
    order.assign((size_t)ntx, 0);
    std::vector<int> pos(off.begin(), off.end() - 1);
    for (int64_t is = 0; is < ntx; ++is)
        order[pos[lev[is] - 1]++] = (uint32_t)is | (inj[is] ? TX_INJ : 0u);
}

// Injector flags from host source data: p channel from src[0]; velocity
// channels from src[1..3] or vn_src (added into them by vn2v!).
static void injectors_from(adf_solver* s, int which, const void* data) {
    const int64_t ntx = s->ntx, nt = s->dom.nt;
    auto nz_row = [&](const void* base, int64_t is) {
        if (s->es == 8) {
            const double* r = (const double*)base + is * nt;
            for (int64_t t = 0; t < nt; ++t) if (r[t] != 0.0) return true;
        } else {
            const float* r = (const float*)base + is * nt;
            for (int64_t t = 0; t < nt; ++t) if (r[t] != 0.0f) return true;
        }
        return false;
    };
    for (int64_t is = 0; is < ntx; ++is) {
        if (which == 0) s->src_inj_p[is] = nz_row(data, is);
        else if (which >= 1 && which <= 3) s->src_inj_v[which - 1][is] = nz_row(data, is);
        else s->vn_inj[is] = nz_row(data, is);
    }
}

// The lists of the one-launch transceiver step (k_tx_step).  A cell is
// written after the sweep when an injector's stencil covers it (pressure
// cells of ip, velocity cells of iv) or when it is a fix-up cell (with its
// mirror cell).  The chain is every transceiver that injects, is a GF map
// output, or has a stencil corner on a written cell; the others only
// record, from cells nothing in the launch writes, so their records do not
// depend on the order.  Transceivers sharing a written cell (and the
// injectors a fix-up cell reads) form groups; the transceiver order
// (levels) and the fix-up only constrain members of one group, so groups
// run independently, whole groups packed into chain blocks of up to
// TXC_BLK_TX transceivers and TXC_BLK_FIX fix-up cells.  A group with more
// than TXC_MAX transceivers, or more than TXC_MAXLEV levels, keeps the
// per-level launches.
static constexpr int TXC_MAX = 64, TXC_MAX_FIX = 4096, TXC_BLK_TX = 8, TXC_BLK_FIX = 256;
static int build_tx_chain(adf_solver* s, const std::vector<uint8_t>& ip, const std::vector<uint8_t>& iv,
                          const std::vector<uint32_t>& order_p, const std::vector<uint32_t>& order_v,
                          const std::vector<uint32_t>& fix) {
    dfree(s->txc_list);
    dfree(s->txc_hdr);
    s->txc_list = nullptr;
    s->txc_hdr = nullptr;
    s->txc_ok = false;
    s->txc_list_h.clear();
    s->txc_hdr_h.clear();
    s->txs_state = 0;
    const int64_t ntx = s->ntx;
    if ((int)s->lev_p.size() - 1 > TXC_MAXLEV || (int)s->lev_v.size() - 1 > TXC_MAXLEV) return ADF_OK;
    const int nx = (int)s->dom.nx, ny = (int)s->dom.ny, nz = (int)s->dom.nz;
    const int64_t n3 = (int64_t)nx * ny * nz;
    auto corners = [&](int64_t is, int ch, int64_t* out) {
        const int32_t* id = &s->h_idx[(size_t)3 * (is + ntx * ch)];
        const int I = id[0], J = id[1], K = id[2];
        if (!(I >= 1 && I < nx && J >= 1 && J < ny && K >= 1 && K < nz)) return false;
        for (int q = 0; q < 8; ++q)
            out[q] = (int64_t)(I - 1 + (q & 1)) + (int64_t)nx * ((J - 1 + ((q >> 1) & 1)) + (int64_t)ny * (K - 1 + (q >> 2)));
        return true;
    };
    std::vector<int> par((size_t)ntx);
    for (int64_t i = 0; i < ntx; ++i) par[i] = (int)i;
    auto find = [&](int x) {
        while (par[x] != x) x = par[x] = par[par[x]];
        return x;
    };
    auto unite = [&](int a, int b) {
        a = find(a);
        b = find(b);
        if (a != b) par[std::max(a, b)] = std::min(a, b);
    };
    // written cells, keyed channel * n3 + cell (channel 0 = P, 1..3 = Vx..Vz),
    // each with one writer; writers of one cell are united
    std::unordered_map<int64_t, int> owner;
    auto own = [&](int64_t key, int is) {
        auto it = owner.find(key);
        if (it == owner.end()) owner.emplace(key, is);
        else unite(it->second, is);
    };
    int64_t cs[8];
    for (int64_t is = 0; is < ntx; ++is) {
        if (ip[is] && corners(is, 0, cs))
            for (int q = 0; q < 8; ++q) own(cs[q], (int)is);
        if (iv[is])
            for (int ch = 1; ch <= 3; ++ch)
                if (corners(is, ch, cs))
                    for (int q = 0; q < 8; ++q) own(ch * n3 + cs[q], (int)is);
    }
    // a fix-up cell belongs with the injectors of the pressure cells it
    // reads (at least one of its two is injected: the list is built so)
    std::vector<int> fix_of(fix.size(), -1);
    for (size_t x = 0; x < fix.size(); ++x) {
        const int ax = (int)(fix[x] >> 30);
        const int64_t c = fix[x] & 0x3fffffffu;
        const int64_t st = ax == 0 ? 1 : (ax == 1 ? nx : (int64_t)nx * ny);
        for (int64_t pc : {c, c - st}) {
            auto it = owner.find(pc);
            if (it == owner.end()) continue;
            if (fix_of[x] < 0) fix_of[x] = it->second;
            else unite(fix_of[x], it->second);
        }
        if (fix_of[x] < 0) return ADF_OK;  // not expected; keep the level launches
        own((ax + 1) * n3 + c, fix_of[x]);
        const int64_t j = (c / nx) % ny, k = c / ((int64_t)nx * ny);
        if (ax == 1 && j == 1) own(2 * n3 + c - nx, fix_of[x]);                 // Vy mirror
        if (ax == 2 && k == 1) own(3 * n3 + c - (int64_t)nx * ny, fix_of[x]);   // Vz mirror
    }
    std::vector<uint8_t> dep((size_t)ntx, 0);
    for (uint32_t e : order_v)
        if (e & TX_GF_OUT) dep[e & TX_INDEX] = 1;
    for (int64_t is = 0; is < ntx; ++is) {
        if (ip[is] || iv[is]) dep[is] = 1;
        for (int ch = 0; ch <= 3; ++ch)
            if (corners(is, ch, cs))
                for (int q = 0; q < 8; ++q) {
                    auto it = owner.find(ch * n3 + cs[q]);
                    if (it != owner.end()) {
                        dep[is] = 1;
                        unite((int)is, it->second);
                    }
                }
    }
    // group sizes, then whole groups packed into chain blocks in the order
    // of their first member
    std::vector<int> gsize((size_t)ntx, 0), gfix((size_t)ntx, 0), blk((size_t)ntx, -1);
    for (int64_t is = 0; is < ntx; ++is)
        if (dep[is]) gsize[find((int)is)]++;
    for (size_t x = 0; x < fix.size(); ++x) gfix[find(fix_of[x])]++;
    int nb = 0, btx = 0, bfix = 0;
    for (int64_t is = 0; is < ntx; ++is) {
        if (!dep[is]) continue;
        const int r = find((int)is);
        if (blk[r] >= 0) continue;
        if (gsize[r] > TXC_MAX || gfix[r] > TXC_MAX_FIX) return ADF_OK;
        if (nb == 0 || btx + gsize[r] > TXC_BLK_TX || bfix + gfix[r] > TXC_BLK_FIX) {
            ++nb;
            btx = bfix = 0;
        }
        blk[r] = nb - 1;
        btx += gsize[r];
        bfix += gfix[r];
    }
    // per block: pressure entries by level, velocity entries by level, fix-up cells
    std::vector<std::vector<uint32_t>> bp((size_t)nb * TXC_MAXLEV), bv((size_t)nb * TXC_MAXLEV), bf((size_t)nb);
    const int nlp = (int)s->lev_p.size() - 1, nlv = (int)s->lev_v.size() - 1;
    for (int L = 0; L < nlp; ++L)
        for (int x = s->lev_p[L]; x < s->lev_p[L + 1]; ++x)
            if (dep[order_p[x] & TX_INDEX]) bp[(size_t)blk[find((int)(order_p[x] & TX_INDEX))] * TXC_MAXLEV + L].push_back(order_p[x]);
    for (int L = 0; L < nlv; ++L)
        for (int x = s->lev_v[L]; x < s->lev_v[L + 1]; ++x)
            if (dep[order_v[x] & TX_INDEX]) bv[(size_t)blk[find((int)(order_v[x] & TX_INDEX))] * TXC_MAXLEV + L].push_back(order_v[x]);
    for (size_t x = 0; x < fix.size(); ++x) bf[blk[find(fix_of[x])]].push_back(fix[x]);
    std::vector<uint32_t> list;
    std::vector<int> hdr((size_t)nb * TXC_HDR, 0);
    for (int b = 0; b < nb; ++b) {
        int* h = &hdr[(size_t)b * TXC_HDR];
        for (int L = 0; L <= TXC_MAXLEV; ++L) {
            h[L] = (int)list.size();
            if (L < TXC_MAXLEV) list.insert(list.end(), bp[(size_t)b * TXC_MAXLEV + L].begin(), bp[(size_t)b * TXC_MAXLEV + L].end());
        }
        for (int L = 0; L <= TXC_MAXLEV; ++L) {
            h[TXC_MAXLEV + 1 + L] = (int)list.size();
            if (L < TXC_MAXLEV) list.insert(list.end(), bv[(size_t)b * TXC_MAXLEV + L].begin(), bv[(size_t)b * TXC_MAXLEV + L].end());
        }
        h[2 * (TXC_MAXLEV + 1)] = (int)list.size();
        list.insert(list.end(), bf[b].begin(), bf[b].end());
        h[2 * (TXC_MAXLEV + 1) + 1] = (int)list.size();
    }
    const int nlist = (int)list.size();
    for (int64_t is = 0; is < ntx; ++is)
        if (!dep[is]) list.push_back((uint32_t)is);
    s->txc_nchain = nb;
    s->txc_nrec = (int)list.size() - nlist;
    s->txc_rec_off = nlist;
    if (!list.empty()) {
        int rc = dmalloc((void**)&s->txc_list, list.size() * 4, "transceiver chain");
        if (rc) return rc;
        HIPCHK(hipMemcpy(s->txc_list, list.data(), list.size() * 4, hipMemcpyHostToDevice));
    }
    if (nb) {
        int rc = dmalloc((void**)&s->txc_hdr, hdr.size() * 4, "transceiver chain");
        if (rc) return rc;
        HIPCHK(hipMemcpy(s->txc_hdr, hdr.data(), hdr.size() * 4, hipMemcpyHostToDevice));
    }
    s->txc_list_h = std::move(list);
    s->txc_hdr_h = std::move(hdr);
    s->txc_ok = true;
    return ADF_OK;
}

// Levels, fixup cells and the mirror decision from the current injector
// flags; uploads the schedule.
int rebuild_schedule(adf_solver* s) {
    const int64_t ntx = s->ntx;
    if (!ntx) return ADF_OK;
    const int nx = (int)s->dom.nx, ny = (int)s->dom.ny, nz = (int)s->dom.nz;
    hipStream_t st = s->ctx->stream;
    std::vector<uint8_t> ip(ntx), iv(ntx);
    for (int64_t is = 0; is < ntx; ++is) {
        ip[is] = s->src_inj_p[is] || s->gf_inj[is];
        iv[is] = s->src_inj_v[0][is] || s->src_inj_v[1][is] || s->src_inj_v[2][is] || s->vn_inj[is] || s->gf_inj[is];
    }
    std::vector<uint32_t> order;
    const int cp[1] = {0};
    make_levels(ntx, nx, ny, nz, s->h_idx, ip, cp, 1, order, s->lev_p);
    HIPCHK(hipMemcpyAsync(s->order_p, order.data(), (size_t)ntx * 4, hipMemcpyHostToDevice, st));
    s->h_order_p = order;
    std::vector<uint32_t> order_v;
    const int cv[3] = {1, 2, 3};
    make_levels(ntx, nx, ny, nz, s->h_idx, iv, cv, 3, order_v, s->lev_v);
    // GF map with write-back: its outputs' vn2v! runs after the
    // extrapolation, and k_tx_v writes its inputs' history (first/next
    // input slot per transceiver)
    dfree(s->gf_slot_first);
    dfree(s->gf_slot_next);
    s->gf_slot_first = s->gf_slot_next = nullptr;
    if (s->gf_kind && s->gf_wb) {
        std::vector<uint8_t> out(ntx, 0);
        for (int64_t i : s->gf_isrc_h) out[i - 1] = 1;
        for (uint32_t& e : order_v)
            if (out[e & TX_INDEX]) e |= TX_GF_OUT;
        const int64_t nr = (int64_t)s->gf_irec_h.size();
        std::vector<int32_t> first(ntx, -1), next(nr, -1);
        for (int64_t r = nr - 1; r >= 0; --r) {
            const int64_t is = s->gf_irec_h[r] - 1;
            next[r] = first[is];
            first[is] = (int32_t)r;
        }
        int rc = dmalloc((void**)&s->gf_slot_first, (size_t)ntx * 4, "GF input slots");
        if (!rc) rc = dmalloc((void**)&s->gf_slot_next, (size_t)nr * 4, "GF input slots");
        if (rc) return rc;
        HIPCHK(hipMemcpy(s->gf_slot_first, first.data(), (size_t)ntx * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(s->gf_slot_next, next.data(), (size_t)nr * 4, hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpyAsync(s->order_v, order_v.data(), (size_t)ntx * 4, hipMemcpyHostToDevice, st));
    s->h_order_v = order_v;

    // Velocity cells whose update reads a pressure cell an injector writes:
    // the fused pass recomputes them after the injection (k_fixup_v).
    std::vector<uint32_t> cells;
    for (int64_t is = 0; is < ntx; ++is) {
        if (!ip[is]) continue;
        const int32_t* id = &s->h_idx[(size_t)3 * is];
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
    HIPCHK(hipStreamSynchronize(st));
    dfree(s->fix_cells);
    s->fix_cells = nullptr;
    s->n_fix = 0;
    s->h_fix.clear();
    if (!cells.empty()) {
        int rc = dmalloc((void**)&s->fix_cells, cells.size() * 4, "fixup list");
        if (rc) return rc;
        HIPCHK(hipMemcpy(s->fix_cells, cells.data(), cells.size() * 4, hipMemcpyHostToDevice));
        s->n_fix = (int)cells.size();
        s->h_fix = cells;
    }

    // The mirrors can ride inside the V update unless a velocity stencil
    // touches Vy planes j<=2 or Vz planes k<=2 (1-based).
    s->fuse_mirror = true;
    for (int64_t is = 0; is < ntx; ++is) {
        for (int ch = 2; ch <= 3; ++ch) {
            const int32_t* id = &s->h_idx[(size_t)3 * (is + ntx * ch)];
            const int I = id[0], J = id[1], K = id[2];
            if (!(I >= 1 && I < nx && J >= 1 && J < ny && K >= 1 && K < nz)) continue;
            if (ch == 2 && J <= 2) s->fuse_mirror = false;
            if (ch == 3 && K <= 2) s->fuse_mirror = false;
        }
    }
    return build_tx_chain(s, ip, iv, order, order_v, cells);
}

int solver_set_txs(adf_solver* s, int64_t ntx, const int64_t* idx, const void* w, const void* src,
                   const void* dir, const void* vn_src) {
    if (s->gf_kind && s->gf_max_index > ntx)
        return fail(ADF_EINVAL, "adf_solver_set_transceivers: the GF map's iRec/iSrc reach transceiver " +
                                    std::to_string(s->gf_max_index) + " of " + std::to_string(ntx) +
                                    "; remove the map (adf_solver_set_gf_map kind 0) or set a new one first");
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

    HIPCHK(hipStreamSynchronize(st));
    s->h_idx = id32;
    s->src_inj_p.assign(ntx, 0);
    for (int c = 0; c < 3; ++c) s->src_inj_v[c].assign(ntx, 0);
    s->vn_inj.assign(ntx, 0);
    s->gf_inj.assign(ntx, 0);
    s->gf_hist_valid = 0;  // the record traces start over
    s->gf_B = -1;
    // a GF map set before keeps writing its sources: they stay injectors
    if (s->gf_kind && s->gf_wb)
        for (int64_t i : s->gf_isrc_h) s->gf_inj[i - 1] = 1;
    const size_t chb = (size_t)ntx * nt * es;
    for (int c = 0; c < 4; ++c) injectors_from(s, c, (const char*)src + c * chb);
    if (vn_src) injectors_from(s, 4, vn_src);
    if ((rc = rebuild_schedule(s))) return rc;
    return ADF_OK;
}

// A source buffer was overwritten from the host (set_trace): refresh the
// injector flags of that buffer and the schedule.
int solver_trace_written(adf_solver* s, int which, const void* data) {
    if (which == 0 || which == 4) {  // p_rec / vn_rec: the GF history is rebuilt from them
        s->gf_hist_valid = 0;
        s->gf_B = -1;
        return ADF_OK;
    }
    if (which >= 5 && which <= 8) injectors_from(s, which - 5, data);
    else if (which == 9) injectors_from(s, 4, data);
    else return ADF_OK;
    return rebuild_schedule(s);
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
    // a previous map with write-back flagged its outputs in the schedule
    const bool had_wb = s->gf_kind && s->gf_wb;
    free_gf(s);
    if (had_wb) {
        const int rc = rebuild_schedule(s);
        if (rc) return rc;
    }
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
    if (writeback) {
        // each output owns its transceiver's source slots (the outputs are
        // written in parallel, then converted by vn2v! once)
        std::vector<uint8_t> seen(s->ntx, 0);
        for (int64_t i = 0; i < n_isrc; ++i) {
            if (seen[iSrc[i] - 1])
                return fail(ADF_EINVAL, "adf_solver_set_gf_map: iSrc names transceiver " + std::to_string(iSrc[i]) +
                                            " twice; with write-back every output needs its own transceiver");
            seen[iSrc[i] - 1] = 1;
        }
    }
    const int nten = kind == 2 ? 4 : 2;
    const size_t n = (size_t)nt_gf * dim2 * dim3;
    int rc;
    // Two device copies of each tensor in the storage type: the far pass's
    // tile-major map (gf_tile_index; zero past nt_gf and the last output)
    // and the near pass's lags 0..GF_M-1 as [s][r][GF_M] (zero past nt_gf).
    const int n_og = (int)((n_isrc + 15) / 16), nkb = (int)((nt_gf + 15) / 16);
    const size_t ntile = (size_t)n_irec * n_og * nkb * 256, nn = (size_t)n_isrc * n_irec * GF_M;
    auto upload = [&](auto* g, void* dtile, void* dnear) -> int {
        using E = std::remove_const_t<std::remove_pointer_t<decltype(g)>>;
        std::vector<E> c(ntile, E(0));
        for (int64_t ir = 0; ir < n_irec; ++ir)
            for (int64_t is = 0; is < n_isrc; ++is) {
                const E* gs = g + gf_base(kind, is, ir, nt_gf, dim2);
                for (int64_t k = 0; k < nt_gf; ++k)
                    c[gf_tile_index(ir, is / 16, k / 16, n_og, nkb, (int)(k % 16) / 4, (int)(is % 16), (int)(k % 4))] =
                        gs[k];
            }
        HIPCHK(hipMemcpy(dtile, c.data(), ntile * sizeof(E), hipMemcpyHostToDevice));
        std::vector<E> d(nn, E(0));
        for (int64_t is = 0; is < n_isrc; ++is)
            for (int64_t ir = 0; ir < n_irec; ++ir) {
                const E* gs = g + gf_base(kind, is, ir, nt_gf, dim2);
                for (int64_t k = 0; k < std::min<int64_t>(nt_gf, GF_M); ++k)
                    d[((size_t)is * n_irec + ir) * GF_M + k] = gs[k];
            }
        HIPCHK(hipMemcpy(dnear, d.data(), nn * sizeof(E), hipMemcpyHostToDevice));
        return ADF_OK;
    };
    for (int t = 0; t < nten; ++t) {
        if (!tensors[t]) return fail(ADF_EINVAL, "null GF tensor");
        if ((rc = dmalloc(&s->gf_t[t], ntile * s->es, "GF tensor (far-pass tiles)"))) return rc;
        if ((rc = dmalloc(&s->gf_tn[t], nn * s->es, "GF tensor (near lags)"))) return rc;
        if (s->es == 8) {
            if ((rc = upload(tensors[t], s->gf_t[t], s->gf_tn[t]))) return rc;
        } else {
            std::vector<float> f(n);
            to_type(tensors[t], f.data(), n);
            if ((rc = upload(f.data(), s->gf_t[t], s->gf_tn[t]))) return rc;
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
    if (writeback)
        for (int64_t i = 0; i < n_isrc; ++i) s->gf_inj[iSrc[i] - 1] = 1;
    s->gf_ng = nt_gf;
    s->gf_d2 = dim2;
    s->gf_d3 = dim3;
    s->gf_nr = n_irec;
    s->gf_ns = n_isrc;
    s->gf_rc = gf_far_rc(n_irec, n_isrc);
    s->gf_nchunk = (int)((n_irec + s->gf_rc - 1) / s->gf_rc);
    s->gf_isrc_h.assign(iSrc, iSrc + n_isrc);
    s->gf_irec_h.assign(iRec, iRec + n_irec);
    s->gf_max_index = 0;
    for (int64_t i = 0; i < n_irec; ++i) s->gf_max_index = std::max(s->gf_max_index, iRec[i]);
    for (int64_t i = 0; i < n_isrc; ++i) s->gf_max_index = std::max(s->gf_max_index, iSrc[i]);
    if ((rc = dmalloc((void**)&s->gf_part, (size_t)s->gf_nchunk * GF_M * 2 * n_isrc * 8, "GF partial sums")))
        return rc;
    s->gf_B = -1;
    s->gf_hist_valid = 0;
    return writeback ? rebuild_schedule(s) : ADF_OK;
}

void* cur_field(adf_solver* s) { return s->fieldb[s->cur]; }
void* cur_mv(adf_solver* s, int a) { return s->mvb[s->cur][a]; }
void* cur_mp(adf_solver* s, int a) { return s->mpb[s->cur][a]; }

int solver_set_option(adf_solver* s, int option, int64_t value) {
    if (option == ADF_OPT_PATH) {
        if (value < 0 || value > 2) return fail(ADF_EINVAL, "ADF_OPT_PATH must be 0, 1 or 2");
        if (value == 2 && !fused_supported(s))
            return fail(ADF_EINVAL, "fused path needs 4 nx ny nz sizeof(T) < 4 GiB, at most 63 row segments and no "
                                    "lane holding cells of both x CPML slabs");
        s->path = (int)value;
        touch(s);
    } else if (option == ADF_OPT_TILE_Y) {
        if (value != 0 && value != SWEEP_TY)
            return fail(ADF_EINVAL, "ADF_OPT_TILE_Y must be 0 or " + std::to_string(SWEEP_TY));
        s->tile_y = (int)value;
        touch(s);
    } else if (option == ADF_OPT_TIME_SWEEP) {
        if (value < 0 || value > (1 << 30)) return fail(ADF_EINVAL, "ADF_OPT_TIME_SWEEP must be >= 0");
        s->time_sweep = (int)value;
    } else if (option == ADF_OPT_BLOCK_Z) {
        if (value < 0 || value > (1 << 20)) return fail(ADF_EINVAL, "ADF_OPT_BLOCK_Z out of range");
        s->block_z = (int)value;
        touch(s);
    } else if (option == ADF_OPT_BLOCK_Z_GEN) {
        if (value < 0 || value > 65535) return fail(ADF_EINVAL, "ADF_OPT_BLOCK_Z_GEN out of range");
        s->block_z_gen = (int)value;
        touch(s);
    } else if (option == ADF_OPT_WAVES) {
        if (value != 0 && value != 1 && value != 2 && value != 4) return fail(ADF_EINVAL, "ADF_OPT_WAVES must be 0, 1, 2 or 4");
        s->waves = (int)value;
        touch(s);
    } else if (option == ADF_OPT_STORE_NT) {
        if (value < -1 || value > 1) return fail(ADF_EINVAL, "ADF_OPT_STORE_NT must be -1, 0 or 1");
        s->store_nt = (int)value;
    } else if (option == ADF_OPT_PIPELINE) {
        if (value != 0 && value != 1) return fail(ADF_EINVAL, "ADF_OPT_PIPELINE must be 0 or 1");
        s->pipeline = (int)value;
        touch(s);
    } else if (option == ADF_OPT_CELLS_PER_LANE) {
        if (value != 0 && value != 1 && value != 2 && value != 4)
            return fail(ADF_EINVAL, "ADF_OPT_CELLS_PER_LANE must be 0, 1, 2 or 4");
        s->cpl_opt = (int)value;
        touch(s);
    } else if (option == ADF_OPT_TX_LAUNCH) {
        if (value < 0 || value > 2) return fail(ADF_EINVAL, "ADF_OPT_TX_LAUNCH must be 0, 1 or 2");
        s->tx_launch = (int)value;
    } else if (option == ADF_OPT_SWEEP_LAUNCH) {
        if (value < ADF_SWEEP_ONE_LAUNCH || value > ADF_SWEEP_TWO_LAUNCHES)
            return fail(ADF_EINVAL, "ADF_OPT_SWEEP_LAUNCH must be 0, 1 or 2");
        s->launch = (int)value;
        touch(s);
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

// Row segments of the fused kernel (xlane): a row of up to 64 cpl cells is
// one wave's; a wider one takes the fewest waves of at most 62 cpl owned
// cells each, with the cells shared out evenly (a multiple of cpl each).
static int64_t fused_nseg(int64_t nx, int cpl) { return nx <= 64 * cpl ? 1 : (nx + 62 * cpl - 1) / (62 * cpl); }
static int fused_xseg(int64_t nx, int cpl) {
    const int64_t n = fused_nseg(nx, cpl);
    if (n == 1) return 64 * cpl;
    return (int)(((nx + n - 1) / n + cpl - 1) / cpl * cpl);
}

// True when no lane of cpl cells (lanes start at multiples of cpl, also in
// row segments) holds updated cells of both x slabs of a phase: the fused
// routines address a lane's x-slab memory from the side of its first slab
// cell.  Only rows whose two slabs (nearly) meet have such a lane.
static bool x_slabs_apart(int64_t nx, int64_t p, int cpl) {
    for (int ph = 0; ph < 2; ++ph) {
        const int64_t w = ph ? p + 1 : p, lo = ph ? 1 : 0, hi = ph ? nx : nx - 1;  // V: [1, nx), P: [0, nx-1)
        if (w <= 0) continue;
        for (int64_t i0 = 0; i0 < nx; i0 += cpl) {
            bool left = false, right = false;
            for (int64_t i = i0; i < i0 + cpl && i < nx; ++i) {
                if (i < lo || i >= hi) continue;
                if (i < w) left = true;
                else if (i >= nx - w) right = true;
            }
            if (left && right) return false;
        }
    }
    return true;
}

// Cells per lane of the fused kernel.  It addresses a field buffer (4
// channels) through one buffer descriptor, so it needs 4 n3 sizeof(T) <
// 4 GiB, a tile entry holds at most 63 row segments, and no lane may hold
// cells of both x slabs (x_slabs_apart; such grids take the two-pass
// kernels, whose slab test is per cell).
static int fused_cpl(const adf_solver* s) {
    const int64_t nx = s->dom.nx;
    const int cpl = s->cpl_opt ? s->cpl_opt : (nx > 128 ? 4 : (nx > 64 ? 2 : 1));
    if (fused_nseg(nx, cpl) > 63) return 0;
    if (4 * s->n3 * (int64_t)s->es >= (int64_t(1) << 32)) return 0;
    if (!x_slabs_apart(nx, s->p, cpl)) return 0;
    return cpl;
}

template <typename T>
static Fused<T> make_fused(adf_solver* s, int ty) {
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
    F.psi_n = (int)s->psi_n;
    F.cpml_p = s->p > 0 && (s->has_mask || s->p > 1);
    F.fuse_mirror = s->fuse_mirror ? 1 : 0;
    for (int a = 0; a < 3; ++a) {
        F.off_mv[a] = (int)s->psi_off[a];
        F.off_mp[a] = (int)s->psi_off[3 + a];
    }
    F.ty = ty;
    F.xseg = fused_xseg(s->dom.nx, std::max(fused_cpl(s), 1));
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

bool fused_supported(const adf_solver* s) { return fused_cpl(s) > 0; }

// Tile shape: TY rows per interior tile, KB planes per interior tile (halved
// from 64 until the launch has 16 tiles per CU: short dependency chains
// and even work per CU matter more than the halo plane each tile reloads)
// and KBG planes per side tile.
void pick_tiles(const adf_solver* s, int* ty, int* kb, int* kbg) {
    const int t = sweep_ty_int((int)s->es, fused_cpl(s));
    int k = s->block_z;
    if (!k) {
        const int64_t ny = s->dom.ny, nz = s->dom.nz;
        const int64_t want = 16LL * s->ctx->num_cu;
        const int64_t tiles_y = (ny + t - 1) / t * fused_nseg(s->dom.nx, std::max(fused_cpl(s), 1));
        k = 64;
        while (k > 8 && tiles_y * ((nz + k - 1) / k) < want) k /= 2;
    }
    *ty = t;
    *kb = k;
    *kbg = s->block_z_gen ? s->block_z_gen : 2;
}

static constexpr int KZ = 8;

// The per-step transceiver / extrapolation / mirror launches shared by both
// paths; S points at the buffers holding the new field.
template <typename T>
static void launch_tx_p(adf_solver* s, const Solve<T>& S, const Txs<T>& X, int it0, hipStream_t st) {
    for (size_t L = 0; L + 1 < s->lev_p.size(); ++L) {
        const int n = s->lev_p[L + 1] - s->lev_p[L];
        if (n > 0)
            hipLaunchKernelGGL((k_tx_p<T>), dim3((8 * n + 255) / 256), dim3(256), 0, st, S.P, S.nx, S.ny, S.nz, X,
                               s->order_p + s->lev_p[L], n, it0);
    }
}

// The blocked GF gather of step t (0-based; the history of t is recorded):
// the far pass of t's block (steps B = GF_M * floor(t / GF_M) onward) when
// the partial sums held are another block's, then the near pass.  Blocks
// start at multiples of GF_M whatever step a run starts at, so the sums of
// a step do not depend on how the steps were split into runs.
template <typename T>
static void launch_gf_gather(adf_solver* s, const Txs<T>& X, int t, hipStream_t st) {
    const int nt = (int)s->dom.nt, ng = (int)s->gf_ng;
    const int nin = (int)s->gf_nr, nout = (int)s->gf_ns;
    const T* g[4] = {(const T*)s->gf_t[0], (const T*)s->gf_t[1], (const T*)s->gf_t[2], (const T*)s->gf_t[3]};
    const int64_t B = (int64_t)(t / GF_M) * GF_M;
    if (s->gf_B != B) {
        const int64_t Bend = std::min<int64_t>(B + GF_M, nt - 1);
        const int Mb = (int)(Bend - B), nchunk = s->gf_nchunk, rc = s->gf_rc;
        const int waves = (nout + 15) / 16 * nchunk;
        const dim3 grid((unsigned)((waves + 3) / 4));
        if (s->gf_kind == 2)
            hipLaunchKernelGGL((k_gf_far<2, T>), grid, dim3(256), 0, st, g[0], g[1], g[2], g[3], ng, nin, nout,
                               s->gf_h0, s->gf_h1, nt, (int)B, Mb, rc, nchunk, s->gf_part);
        else if (s->gf_kind == 1)
            hipLaunchKernelGGL((k_gf_far<1, T>), grid, dim3(256), 0, st, g[0], g[1], g[2], g[3], ng, nin, nout,
                               s->gf_h0, s->gf_h1, nt, (int)B, Mb, rc, nchunk, s->gf_part);
        else
            hipLaunchKernelGGL((k_gf_far<3, T>), grid, dim3(256), 0, st, g[0], g[1], g[2], g[3], ng, nin, nout,
                               s->gf_h0, s->gf_h1, nt, (int)B, Mb, rc, nchunk, s->gf_part);
        hipLaunchKernelGGL(k_gf_far_sum, dim3((unsigned)((Mb * 2 * nout + 255) / 256)), dim3(256), 0, st, s->gf_part,
                           nout, Mb, nchunk);
        s->gf_B = B;
    }
    const int m = (int)(t - B);
    const double* part = s->gf_part;
    const T* gn[4] = {(const T*)s->gf_tn[0], (const T*)s->gf_tn[1], (const T*)s->gf_tn[2], (const T*)s->gf_tn[3]};
    if (s->gf_kind == 2)
        hipLaunchKernelGGL((k_gf_near<2, T, T>), dim3((unsigned)nout), dim3(1024), 0, st, X, gn[0], gn[1], gn[2],
                           gn[3], ng, s->gf_isrc, nout, nin, s->gf_h0, s->gf_h1, s->dom.z0, t, m, part);
    else if (s->gf_kind == 1)
        hipLaunchKernelGGL((k_gf_near<1, T, T>), dim3((unsigned)nout), dim3(1024), 0, st, X, gn[0], gn[1], gn[2],
                           gn[3], ng, s->gf_isrc, nout, nin, s->gf_h0, s->gf_h1, s->dom.z0, t, m, part);
    else
        hipLaunchKernelGGL((k_gf_near<3, T, T>), dim3((unsigned)nout), dim3(1024), 0, st, X, gn[0], gn[1], gn[2],
                           gn[3], ng, s->gf_isrc, nout, nin, s->gf_h0, s->gf_h1, s->dom.z0, t, m, part);
}

static GfHist gf_hist(const adf_solver* s) {
    return GfHist{s->gf_slot_first, s->gf_slot_next, s->gf_h0, s->gf_h1, s->gf_kind, s->dom.z0};
}

// tx_done: the velocity transceivers ran in k_tx_step.
template <typename T>
static void launch_after_v(adf_solver* s, const Solve<T>& S, const Txs<T>& X, int64_t it, hipStream_t st,
                           bool tx_done = false) {
    const int it0 = (int)(it - 1);
    const int nt = (int)s->dom.nt;
    if (s->ntx) {
        const GfHist H = gf_hist(s);
        for (size_t L = 0; !tx_done && L + 1 < s->lev_v.size(); ++L) {
            const int n = s->lev_v[L + 1] - s->lev_v[L];
            if (n > 0)
                hipLaunchKernelGGL((k_tx_v<T>), dim3((32 * n + 255) / 256), dim3(256), 0, st, S.Vx, S.Vy, S.Vz, S.nx, S.ny,
                                   S.nz, X, s->order_v + s->lev_v[L], n, it0, H);
        }
        // The extrapolation (extrapolation.jl:101-249) and the map outputs'
        // vn2v!; without write-back nothing reads the map's history.
        if (s->gf_kind && s->gf_wb && it < nt) launch_gf_gather<T>(s, X, it0, st);
    }
    if (!S.fuse_mirror) {
        const int n = std::max(S.nx * S.nz, S.nx * S.ny);
        hipLaunchKernelGGL((k_mirror<T>), dim3((n + 255) / 256), dim3(256), 0, st, S);
    }
}

// Split the grid into tiles of ty rows: interior tiles, whose rows and
// planes (with their halos) avoid every y/z CPML slab, the mirrored planes
// 1,2 and the rigid planes n, in blocks of kb planes, and general tiles (all
// others) in blocks of kbg planes; both lists in (z-block, y-tile) order,
// entries as xcd_tile describes, for member `m`.
// Rows (j, k) whose cells are all fluid with all faces open (every row
// when the solve has no mask), from the device mask.
static int open_rows(adf_solver* s, std::vector<uint8_t>& open) {
    const int64_t nx = s->dom.nx, ny = s->dom.ny, nz = s->dom.nz, nwx = s->nwx;
    open.assign((size_t)(ny * nz), 1);
    if (!s->has_mask) return ADF_OK;
    if (s->open_ok) {
        open = s->open_h;
        return ADF_OK;
    }
    std::vector<uint32_t> bits((size_t)(4 * nwx * ny * nz));
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
    HIPCHK(hipMemcpy(bits.data(), s->mask, bits.size() * 4, hipMemcpyDeviceToHost));
    const uint32_t last = (nx % 32) ? ((1u << (nx % 32)) - 1u) : 0xffffffffu;
    // Bits whose value no routine reads are left out: P at i = nx-1 and Vx
    // at i = 0 are never updated (the sweep keeps their old values
    // whatever the mask); Vy on row j = 0 and Vz on plane k = 0 only take
    // the mirrors of row / plane 1 (general tiles hold them, and their
    // update keeps the old values there).  The mask rule closes Vx at
    // i = 0 (and Vy/Vz on the first row/plane) everywhere, so without this
    // no row of a masked grid would be open.
    const int64_t lastw = (nx - 1) >> 5;
    for (int64_t c = 0; c < 4; ++c)
        for (int64_t r = 0; r < ny * nz; ++r) {
            if ((c == 2 && r % ny == 0) || (c == 3 && r < ny)) continue;
            const uint32_t* w = &bits[(size_t)((c * ny * nz + r) * nwx)];
            for (int64_t q = 0; q < nwx; ++q) {
                uint32_t want = q == nwx - 1 ? last : 0xffffffffu;
                if (c == 0 && q == lastw) want &= ~(1u << ((nx - 1) & 31));
                if (c == 1 && q == 0) want &= ~1u;
                if ((w[q] & want) != want) { open[(size_t)r] = 0; break; }
            }
        }
    s->open_h = open;
    s->open_ok = true;
    return ADF_OK;
}

// Interior and side lists in (z-block, row) order, then both merged into
// LIST_ALL: a launch of W waves per workgroup gives XCD x the x-th of 8
// contiguous chunks of the list (xcd_tile), so chunk x holds its share of
// the side tiles first (their short chains start early) and then its share
// of the interior tiles, each share taken in list order.
static void merge_lists(const std::vector<uint2>& side, const std::vector<uint2>& in, int W, std::vector<uint2>& out) {
    const int64_t S = (int64_t)side.size(), n = S + (int64_t)in.size();
    const int64_t nb = (n + W - 1) / W;
    auto start = [&](int x) {  // first tile of chunk x (xcd_chunk_start over blocks)
        const int64_t b = x * (nb >> 3) + std::min<int64_t>(x, nb & 7);
        return std::min<int64_t>(b * W, n);
    };
    out.clear();
    for (int x = 0; x < 8; ++x) {
        const int64_t a = start(x), b = start(x + 1);
        const int64_t sa = n ? S * a / n : 0, sb = n ? S * b / n : 0;
        out.insert(out.end(), side.begin() + sa, side.begin() + sb);
        out.insert(out.end(), in.begin() + (a - sa), in.begin() + (b - sb));
    }
}

static void split_tiles(const adf_solver* s, int ty, int kb, int kbg, int W, uint32_t m,
                        const std::vector<uint8_t>& open, std::vector<uint2> (&out)[TILE_LISTS]) {
    const int ny = (int)s->dom.ny, nz = (int)s->dom.nz, p = (int)s->p, qv = p + 1;
    const bool cpP = s->p > 0 && (s->has_mask || s->p > 1), cpV = qv > 1;
    const int tiles_y = (ny + ty - 1) / ty;
    const int nseg = (int)fused_nseg(s->dom.nx, std::max(fused_cpl(s), 1));
    kb = std::min(kb, 1023);
    kbg = std::min(kbg, 1023);
    // regular rows / planes: owned ones >= 2 (no mirror), the halo >= 1,
    // and nothing at or past n-1 (rigid plane, edge)
    auto rgen = [&](int t) { const int j0 = t * ty; return j0 < 2 || j0 + ty > ny - 1; };
    auto in_slab = [&](int i, int n, bool on, int w) { return on && (i < w || i >= n - w); };
    auto rslab = [&](int t) {
        const int j0 = t * ty;
        for (int j = j0 - 1; j < j0 + ty; ++j) {
            if (in_slab(j, ny, cpP, p)) return true;
            if (j >= j0 && in_slab(j, ny, cpV, qv)) return true;
        }
        return false;
    };
    auto zslab = [&](int k0, int k1) {
        for (int k = k0 - 1; k < k1; ++k) {
            if (in_slab(k, nz, cpP, p)) return true;
            if (k >= k0 && in_slab(k, nz, cpV, qv)) return true;
        }
        return false;
    };
    // a tile needs the mask if a row of it or of its halo (rows j0-1..j0+ty,
    // planes k0-1..k1) is not fully open
    // entry of the tile of rows [j0, j0+h), planes [k0, k1), row segment seg
    auto tile = [&](int j0, int h, int k0, int k1, uint32_t kind, int seg) {
        bool all_open = true;
        for (int k = std::max(k0 - 1, 0); k <= std::min(k1, nz - 1) && all_open; ++k)
            for (int j = std::max(j0 - 1, 0); j <= std::min(j0 + h, ny - 1); ++j)
                if (!open[(size_t)k * ny + j]) { all_open = false; break; }
        return make_uint2((all_open ? 0u : 0x80000000u) | (kind << 28) | (m << 16) | (uint32_t)j0,
                          ((uint32_t)k0 << 16) | ((uint32_t)seg << 10) | (uint32_t)(k1 - k0));
    };
    // z ranges: general [0, 2), regular [2, nz-1) split at the z slabs
    // (interior planes [zi0, zi1): halo and owned planes clear of both
    // phases' slabs), general [nz-1, nz).  Interior tiles take kb planes,
    // the side tiles (short dependency chains next to the long interior
    // ones) kbg planes.
    std::vector<std::pair<int, int>> zb, zs;
    auto blocks = [&](std::vector<std::pair<int, int>>& out, int a, int b, int w) {
        for (int k = a; k < b; k += w) out.push_back({k, std::min(k + w, b)});
    };
    const int ra = std::min(2, nz), rb = std::max(ra, nz - 1);
    int zi0 = std::max(ra, cpP || cpV ? p + 1 : ra), zi1 = std::min(rb, nz - p - 1);
    if (zi0 >= zi1) zi0 = zi1 = rb;
    blocks(zb, 0, ra, kbg);
    blocks(zb, ra, zi0, kbg);
    blocks(zb, zi0, zi1, kb);
    blocks(zb, zi1, rb, kbg);
    blocks(zb, rb, nz, kbg);
    blocks(zs, 0, nz, kbg);
    // rows handled by the general routine throughout take one plane per
    // tile: that routine waits out a memory round trip per plane, so its
    // chains are kept short
    std::vector<std::pair<int, int>> z1;
    blocks(z1, 0, nz, 1);
    // (k0, t, entry) in (z-block, y-tile) order, so neighbouring entries
    // share halo rows
    // Rows are classified in tiles of ty rows; side tiles are then cut into
    // tiles of sweep_ty_side rows (a regular tile's row subsets are regular,
    // and the slab routines are exact for rows outside the slab).
    std::vector<std::tuple<int, int, uint2>> lists[TILE_LISTS];
    const int tys = std::min(ty, sweep_ty_side(fused_cpl(s)));
    for (int t = 0; t < tiles_y; ++t) {
        const bool yg = rgen(t), ys = !yg && rslab(t);
        const int j0 = t * ty, j1 = std::min(j0 + ty, ny);
        for (auto& z : yg ? z1 : (ys ? zs : zb)) {
            uint32_t kind = TILE_GEN;
            if (!yg && z.first >= 2 && z.second <= nz - 1) {
                const bool zl = zslab(z.first, z.second);
                kind = ys ? (zl ? TILE_SYZ : TILE_SY) : (zl ? TILE_SZ : TILE_INT);
            }
            if (kind == TILE_INT) {
                // the interior list holds unmasked interior tiles only (its
                // launch may compile their routine alone); masked interior
                // tiles go with the side tiles
                for (int g = 0; g < nseg; ++g) {
                    const uint2 e = tile(j0, ty, z.first, z.second, kind, g);
                    lists[(e.x >> 31) ? LIST_SIDE : LIST_INT].emplace_back(z.first, j0, e);
                }
                continue;
            }
            for (int j = j0; j < j1; j += tys)
                for (int g = 0; g < nseg; ++g)
                    lists[LIST_SIDE].emplace_back(z.first, j, tile(j, tys, z.first, z.second, kind, g));
        }
    }
    auto order = [](const std::tuple<int, int, uint2>& a, const std::tuple<int, int, uint2>& b) {
        return std::get<0>(a) != std::get<0>(b) ? std::get<0>(a) < std::get<0>(b) : std::get<1>(a) < std::get<1>(b);
    };
    for (int c : {LIST_INT, LIST_SIDE}) {
        std::stable_sort(lists[c].begin(), lists[c].end(), order);  // segments of a row stay adjacent
        out[c].clear();
        for (auto& e : lists[c]) out[c].push_back(std::get<2>(e));
    }
    merge_lists(out[LIST_SIDE], out[LIST_INT], W, out[LIST_ALL]);
}

static int upload_tiles(const std::vector<uint2>& v, uint2** dst, int* n) {
    dfree(*dst);
    *dst = nullptr;
    *n = (int)v.size();
    if (v.empty()) return ADF_OK;
    int rc = dmalloc((void**)dst, v.size() * sizeof(uint2), "tile list");
    if (rc) return rc;
    HIPCHK(hipMemcpy(*dst, v.data(), v.size() * sizeof(uint2), hipMemcpyHostToDevice));
    return ADF_OK;
}

static int check_tiling(const adf_solver* s, int ty) {
    (void)ty;
    if (s->dom.ny > 65535 || s->dom.nz > 65535)
        return fail(ADF_EINVAL, "grid too large for the fused sweep's tile list (ny and nz must be < 65536)");
    return ADF_OK;
}

static int prepare_tiles(adf_solver* s, int ty, int kb, int kbg) {
    const int W = s->waves ? s->waves : 1;
    if (s->tl_ty == ty && s->tl_kb == kb && s->tl_kbg == kbg && s->tl_w == W) return ADF_OK;
    int rc = check_tiling(s, ty);
    if (rc) return rc;
    std::vector<uint8_t> open;
    if ((rc = open_rows(s, open))) return rc;
    std::vector<uint2> lists[TILE_LISTS];
    split_tiles(s, ty, kb, kbg, W, 0, open, lists);
    for (int c = 0; c < TILE_LISTS; ++c)
        if ((rc = upload_tiles(lists[c], &s->tiles[c], &s->ntl[c]))) return rc;
    s->tiles_all_h = lists[LIST_ALL];
    s->txs_state = 0;
    if (!s->ev_fork) {
        HIPCHK(hipEventCreateWithFlags(&s->ev_fork, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&s->ev_join, hipEventDisableTiming));
    }
    s->tl_ty = ty;
    s->tl_kb = kb;
    s->tl_kbg = kbg;
    s->tl_w = W;
    return ADF_OK;
}

// The tile order and wait lists of transceivers carried by the sweep launch
// (ADF_OPT_TX_LAUNCH = 2) for the current all-tiles list (tiles of ty
// interior rows, one wave per workgroup) and chain.  A chain block waits
// for the owner tiles of every cell its groups touch: the stencil corners
// of its transceivers on all four channels, its fix-up cells with the two
// pressure cells each reads and its mirror cell, and for a cell on row or
// plane 0 also the owner of row or plane 1 (which writes the mirrored
// velocities there).  Those tiles are marked TILE_SIGNAL and moved to the
// front of their XCD's chunk (stable), so they are stored early in the
// launch.  Leaves txs_state -1 when the lists do not have the expected
// shape or when more than TXS_MAX_SIGNAL tiles would signal (the step then
// keeps k_tx_step).
static constexpr int TXS_MAX_SIGNAL = 64;
static int build_txs(adf_solver* s, int ty) {
    dfree(s->txs_tiles); dfree(s->txs_wait); dfree(s->txs_wait_off); dfree(s->txs_flag);
    s->txs_tiles = nullptr;
    s->txs_wait = s->txs_wait_off = s->txs_flag = nullptr;
    s->txs_state = -1;
    const std::vector<uint2>& L = s->tiles_all_h;
    const int n = (int)L.size(), nb = s->txc_nchain;
    if (!n || (int)s->txc_hdr_h.size() != nb * TXC_HDR) return ADF_OK;
    const int64_t nx = s->dom.nx, ny = s->dom.ny, nz = s->dom.nz, ntx = s->ntx;
    const int cpl = std::max(fused_cpl(s), 1);
    const int nseg = (int)fused_nseg(nx, cpl), xseg = fused_xseg(nx, cpl);
    const int tys = std::min(ty, sweep_ty_side(fused_cpl(s)));
    // owner tile of each (plane, row, segment)
    std::vector<int> own((size_t)(ny * nz * nseg), -1);
    for (int t = 0; t < n; ++t) {
        const uint2 e = L[t];
        if (e.x & TILE_SIGNAL) return ADF_OK;
        const int j0 = (int)(e.x & 0xffffu), kind = (int)((e.x >> 28) & 7u);
        const int k0 = (int)(e.y >> 16), k1 = k0 + (int)(e.y & 0x3ffu), seg = (int)((e.y >> 10) & 0x3fu);
        const int h = kind == TILE_INT ? ty : tys;
        if (seg >= nseg || k1 > nz) return ADF_OK;
        for (int k = k0; k < k1; ++k)
            for (int j = j0; j < std::min<int64_t>(j0 + h, ny); ++j) {
                int& o = own[((size_t)k * ny + j) * nseg + seg];
                if (o >= 0) return ADF_OK;
                o = t;
            }
    }
    for (int o : own)
        if (o < 0) return ADF_OK;
    auto owner = [&](int64_t i, int64_t j, int64_t k) {
        return own[((size_t)k * ny + j) * nseg + std::min<int64_t>(i / xseg, nseg - 1)];
    };
    std::vector<std::vector<int>> wl((size_t)nb);
    std::vector<uint8_t> sig((size_t)n, 0);
    bool bad = false;  // a cell outside the grid (not expected): keep k_tx_step
    for (int b = 0; b < nb; ++b) {
        std::vector<int>& w = wl[b];
        auto add = [&](int64_t c) {
            if (c < 0 || c >= s->n3) {
                bad = true;
                return;
            }
            const int64_t i = c % nx, j = (c / nx) % ny, k = c / (nx * ny);
            w.push_back(owner(i, j, k));
            if (j == 0 && ny > 1) w.push_back(owner(i, 1, k));
            if (k == 0 && nz > 1) w.push_back(owner(i, j, 1));
        };
        auto add_tx = [&](int64_t is) {
            for (int ch = 0; ch < 4; ++ch) {
                const int32_t* id = &s->h_idx[(size_t)3 * (is + ntx * ch)];
                const int64_t I = id[0], J = id[1], K = id[2];
                if (!(I >= 1 && I < nx && J >= 1 && J < ny && K >= 1 && K < nz)) continue;
                for (int q = 0; q < 8; ++q)
                    add((I - 1 + (q & 1)) + nx * ((J - 1 + ((q >> 1) & 1)) + ny * (K - 1 + (q >> 2))));
            }
        };
        const int* h = &s->txc_hdr_h[(size_t)b * TXC_HDR];
        for (int x = h[0]; x < h[TXC_MAXLEV]; ++x) add_tx(s->txc_list_h[x] & TX_INDEX);
        for (int x = h[TXC_MAXLEV + 1]; x < h[2 * TXC_MAXLEV + 1]; ++x) add_tx(s->txc_list_h[x] & TX_INDEX);
        for (int x = h[2 * (TXC_MAXLEV + 1)]; x < h[2 * (TXC_MAXLEV + 1) + 1]; ++x) {
            const uint32_t e = s->txc_list_h[x];
            const int ax = (int)(e >> 30);
            const int64_t c = e & 0x3fffffffu;
            const int64_t st = ax == 0 ? 1 : (ax == 1 ? nx : nx * ny);
            const int64_t j = (c / nx) % ny, k = c / (nx * ny);
            add(c);
            add(c - st);
            if (ax == 1 && j == 1) add(c - nx);
            if (ax == 2 && k == 1) add(c - nx * ny);
        }
        std::sort(w.begin(), w.end());
        w.erase(std::unique(w.begin(), w.end()), w.end());
        for (int t : w) sig[t] = 1;
    }
    // a signal tile's agent-scope release writes back its XCD's L2: worth it
    // for the few tiles of a point source's chain, not for a surface's
    if (bad || std::count(sig.begin(), sig.end(), (uint8_t)1) > TXS_MAX_SIGNAL) return ADF_OK;
    // signal tiles first within each XCD chunk (xcd_tile's chunks of n)
    std::vector<int> pos((size_t)n);
    std::vector<uint2> out;
    out.reserve(n);
    for (int x = 0; x < 8; ++x) {
        const int a = x * (n >> 3) + std::min(x, n & 7), z = a + (n >> 3) + (x < (n & 7) ? 1 : 0);
        for (int pass = 1; pass >= 0; --pass)
            for (int t = a; t < z; ++t)
                if (sig[t] == pass) {
                    pos[t] = (int)out.size();
                    out.push_back(pass ? make_uint2(L[t].x | TILE_SIGNAL, L[t].y) : L[t]);
                }
    }
    std::vector<int> off(1, 0), all;
    for (int b = 0; b < nb; ++b) {
        for (int t : wl[b]) all.push_back(pos[t]);
        off.push_back((int)all.size());
    }
    if (all.empty()) all.push_back(0);
    int rc;
    if ((rc = dmalloc((void**)&s->txs_tiles, out.size() * sizeof(uint2), "transceiver tiles")) ||
        (rc = dmalloc((void**)&s->txs_wait_off, off.size() * 4, "transceiver tiles")) ||
        (rc = dmalloc((void**)&s->txs_wait, all.size() * 4, "transceiver tiles")) ||
        (rc = dmalloc((void**)&s->txs_flag, (size_t)n * 4, "transceiver tiles")))
        return rc;
    if (!s->txs_err) {
        if ((rc = dmalloc((void**)&s->txs_err, 4, "transceiver tiles"))) return rc;
        HIPCHK(hipMemset(s->txs_err, 0, 4));
    }
    if (!s->txs_err_h) {
        HIPCHK(hipHostMalloc((void**)&s->txs_err_h, 4, hipHostMallocDefault));
        *s->txs_err_h = 0;
    }
    HIPCHK(hipMemcpy(s->txs_tiles, out.data(), out.size() * sizeof(uint2), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(s->txs_wait_off, off.data(), off.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(s->txs_wait, all.data(), all.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(s->txs_flag, 0, (size_t)n * 4));
    s->txs_state = 1;
    return ADF_OK;
}

// Start/stop events around one field-update launch (ADF_OPT_TIME_SWEEP) of
// a sampled step.  The events only time: they skip the system-scope fence
// (cache write-back and invalidation) that a default event adds after the
// launch, which cost the following step ~18 us at 256^3.
static void sweep_mark(adf_solver* s, hipStream_t st, bool sampled) {
    if (!sampled) return;
    if ((int)s->sweep_ev.size() <= s->sweep_used) {
        hipEvent_t e;
        if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return;
        s->sweep_ev.push_back(e);
    }
    (void)hipEventRecord(s->sweep_ev[s->sweep_used++], st);
}

static TxChain tx_chain_of(const adf_solver* s) {
    TxChain C{};
    C.list = s->txc_list;
    C.hdr = s->txc_hdr;
    C.rec = s->txc_list + s->txc_rec_off;
    C.nchain = s->txc_nchain;
    C.n_rec = s->txc_nrec;
    C.nlp = (int)s->lev_p.size() - 1;
    C.nlv = (int)s->lev_v.size() - 1;
    return C;
}

template <typename T>
static TxSweep<T> tx_sweep_of(adf_solver* s, const Txs<T>& X) {
    TxSweep<T> S{};
    S.X = X;
    S.C = tx_chain_of(s);
    S.wait_off = s->txs_wait_off;
    S.wait = s->txs_wait;
    S.tflag = s->txs_flag;
    S.err = s->txs_err;
    S.nrb = (S.C.n_rec + TXS_REC - 1) / TXS_REC;
    return S;
}

// One time step it (1-based) of forward_onestep! on stream st.  txs: the
// step's transceivers ride in the sweep launch (ADF_OPT_TX_LAUNCH = 2), the
// records of the previous step with them.
template <typename T, bool EXACT>
static void step_impl(adf_solver* s, int64_t it, hipStream_t st, bool fused, int ty, bool txs) {
    const int it0 = (int)(it - 1);
    Txs<T> X = make_txs<T>(s);
    const bool mark = s->time_sweep > 0 && s->sweep_seq % s->time_sweep == 0;
    ++s->sweep_seq;
    if (fused) {
        Fused<T> F = make_fused<T>(s, ty);
        TxSweep<T> XS{};
        if (txs) {
            XS = tx_sweep_of<T>(s, X);
            s->txs_epoch = s->txs_epoch % 0x7ffffffe + 1;  // never 0 (the flags' initial value)
            XS.epoch = s->txs_epoch;
            XS.nbt = s->ntl[LIST_ALL];
            if (!s->txs_pending) XS.nrb = 0;
            XS.it_rec = it0 - 1;
            XS.it0 = it0;
        }
        sweep_mark(s, st, mark);
        launch_fused<T, EXACT>(s, F, fused_cpl(s), st, txs ? &XS : nullptr);
        sweep_mark(s, st, mark);
        s->cur ^= 1;  // the new buffers become current
        Solve<T> S = make_solve<T>(s);
        const bool one = s->ntx && s->txc_ok && s->tx_launch;
        if (txs) {
            s->txs_pending = true;
        } else if (one) {
            // every transceiver of the step in one launch (F still holds
            // old = 0 and new = 1 buffers)
            const TxChain C = tx_chain_of(s);
            const int nblk = C.nchain + (C.n_rec + 7) / 8;
            if (nblk)
                hipLaunchKernelGGL((k_tx_step<T, EXACT>), dim3(nblk), dim3(256), 0, st, F, X, C, it0, gf_hist(s));
        } else if (s->ntx) {
            launch_tx_p<T>(s, S, X, it0, st);
            if (s->n_fix) {
                // recompute the velocity cells that read injected pressure
                // (F still holds old = 0 and new = 1 buffers)
                hipLaunchKernelGGL((k_fixup_v<T, EXACT>), dim3((s->n_fix + 255) / 256), dim3(256), 0, st, F,
                                   s->fix_cells, s->n_fix);
            }
        }
        launch_after_v<T>(s, S, X, it, st, one);
    } else {
        Solve<T> S = make_solve<T>(s);
        const dim3 blk(64, 4, 1);
        const int kz = s->block_z ? s->block_z : KZ;
        const dim3 gp((S.nx - 1 + 63) / 64, (S.ny + 3) / 4, (S.nz + kz - 1) / kz);
        const dim3 gv((S.nx + 63) / 64, (S.ny + 3) / 4, (S.nz + kz - 1) / kz);
        sweep_mark(s, st, mark);
        hipLaunchKernelGGL((k_update_p<T, EXACT>), gp, blk, 0, st, S, kz);
        sweep_mark(s, st, mark);
        if (s->ntx) launch_tx_p<T>(s, S, X, it0, st);
        sweep_mark(s, st, mark);
        hipLaunchKernelGGL((k_update_v<T, EXACT>), gv, blk, 0, st, S, kz);
        sweep_mark(s, st, mark);
        launch_after_v<T>(s, S, X, it, st);
    }
}

template <typename T, bool EXACT>
static int run_impl(adf_solver* s, int64_t start, int64_t end) {
    hipStream_t st = s->ctx->stream;
    if (s->path == 2 && !fused_supported(s))  // cells per lane changed after the fused path was chosen
        return fail(ADF_EINVAL, "fused path not available for this grid at these cells per lane");
    const bool fused = s->path == 2 || (s->path == 0 && fused_supported(s));
    int ty = 2, kb = 16, kbg = 8;
    bool txs = false;
    if (fused) {
        pick_tiles(s, &ty, &kb, &kbg);
        int rc = prepare_tiles(s, ty, kb, kbg);
        if (rc) return rc;
        if (s->tx_launch == 2 && s->ntx && s->txc_ok && s->launch == ADF_SWEEP_ONE_LAUNCH &&
            (s->waves ? s->waves : 1) == 1 && !s->gf_kind) {
            if (s->txs_state == 0 && (rc = build_txs(s, ty))) return rc;
            txs = s->txs_state == 1;
        }
    }
    s->txs_pending = false;
    s->sweep_used = 0;
    s->sweep_avg_ms = 0.0;
    s->sweep_n = 0;
    s->sweep_seq = 0;
    if (s->gf_kind && s->gf_wb && s->gf_hist_valid < start - 1) {
        // the extrapolation reads the record traces of every earlier step
        // (extrapolation.jl:202-249): steps this solver has not run, or
        // whose traces were replaced, are taken from the traces
        const int t0 = (int)s->gf_hist_valid, t1 = (int)(start - 1);
        const int64_t n = s->gf_nr * (int64_t)(t1 - t0);
        Txs<T> X = make_txs<T>(s);
        hipLaunchKernelGGL((k_gf_hist_fill<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, X, s->gf_irec,
                           (int)s->gf_nr, s->gf_kind, s->dom.z0, s->gf_h0, s->gf_h1, t0, t1);
        s->gf_B = -1;
    }
    HIPCHK(hipEventRecord(s->ev0, st));
    for (int64_t it = start; it <= end; ++it) {
        step_impl<T, EXACT>(s, it, st, fused, ty, txs);
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
    if (s->txs_pending) {
        // the last step's records, from the current buffers
        TxSweep<T> XS = tx_sweep_of<T>(s, make_txs<T>(s));
        XS.it_rec = (int)(end - 1);
        if (XS.nrb)
            hipLaunchKernelGGL((k_tx_records<T>), dim3(XS.nrb), dim3(64), 0, st, make_fused<T>(s, ty), XS);
        s->txs_pending = false;
    }
    // the wait flag comes back with the run's last event (no extra sync)
    if (txs) HIPCHK(hipMemcpyAsync(s->txs_err_h, s->txs_err, 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipGetLastError());
    if (s->gf_kind && s->gf_wb) s->gf_hist_valid = std::max(s->gf_hist_valid, end);
    HIPCHK(hipEventRecord(s->ev1, st));
    HIPCHK(wait_event(s->ev1));
    if (txs && *s->txs_err_h) {
        *s->txs_err_h = 0;
        HIPCHK(hipMemset(s->txs_err, 0, 4));
        return fail(ADF_EDEVICE, "transceiver step: a chain block's wait for its tiles ran out (results invalid)");
    }
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
    int rc;
    if (s->dtype == ADF_F64) rc = run_impl<double, true>(s, start, end);
    else if (s->dtype == ADF_F32_REF) rc = run_impl<float, true>(s, start, end);
    else rc = run_impl<float, false>(s, start, end);
    if (!rc) s->next_it = end + 1;
    return rc;
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
    (void)hipStreamSynchronize(st);
    dfree(buf);
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
    if ((rc = dmalloc((void**)&solid, pl * 4, "mask solid"))) { dfree(dbuf); return rc; }
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
    (void)hipStreamSynchronize(st);
    dfree(dbuf);
    dfree(solid);
    if (e != hipSuccess) return fail(ADF_EDEVICE, std::string("mask build: ") + hipGetErrorString(e));
    return ADF_OK;
}


// ===================================================================== batch

static void batch_drop_tables(adf_batch* b) {
    if (b->gexec) { (void)hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
    if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
    b->gexec_steps = 0;
    dfree(b->fs); dfree(b->td); dfree(b->it0); dfree(b->cur); dfree(b->txp); dfree(b->txv); dfree(b->fix);
    dfree(b->mirror); dfree(b->chains); dfree(b->txmap); dfree(b->txdone);
    b->chains = nullptr;
    b->txmap = nullptr;
    b->txdone = nullptr;
    b->n_txmap = 0;
    b->fs = b->td = nullptr;
    b->it0 = b->cur = b->mirror = nullptr;
    b->txp = b->txv = nullptr;
    b->fix = nullptr;
    for (auto& g : b->groups)
        for (int h = 0; h < TILE_LISTS; ++h) { dfree(g.tiles[h]); dfree(g.ctr[h]); }
    b->groups.clear();
    b->txp_lev.clear();
    b->txv_lev.clear();
    b->n_fix = b->n_mirror = 0;
}

void batch_free(adf_batch* b) {
    for (adf_solver* s : b->members) s->batch = nullptr;
    batch_drop_tables(b);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);

}

int batch_add(adf_batch* b, adf_solver* s) {
    if (s->batch) return fail(ADF_ESTATE, "adf_batch_add: solver already belongs to a batch");
    if (s->ctx != b->ctx) return fail(ADF_EINVAL, "adf_batch_add: solver uses another context");
    if (!b->members.empty() && s->dtype != b->dtype) return fail(ADF_EINVAL, "adf_batch_add: batch members share one dtype");
    if (s->gf_kind) return fail(ADF_EINVAL, "adf_batch_add: solvers with a GF map run alone (adf_solver_run)");
    if (!fused_supported(s))
        return fail(ADF_EINVAL, "adf_batch_add: batched solves need the fused sweep (4 nx ny nz sizeof(T) < 4 GiB, "
                                "at most 63 row segments, no lane holding cells of both x CPML slabs)");
    if ((int64_t)b->members.size() >= 4096) return fail(ADF_EINVAL, "adf_batch_add: at most 4096 members");
    if (s->ntx >= (1 << 19)) return fail(ADF_EINVAL, "adf_batch_add: at most 524287 transceivers per member");
    if (b->members.empty()) b->dtype = s->dtype;
    b->members.push_back(s);
    s->batch = b;
    b->dirty = true;
    return ADF_OK;
}

int batch_remove(adf_batch* b, adf_solver* s) {
    auto it = std::find(b->members.begin(), b->members.end(), s);
    if (it == b->members.end()) return fail(ADF_EINVAL, "adf_batch_remove: not a member");
    (void)hipStreamSynchronize(b->ctx->stream);
    b->members.erase(it);
    s->batch = nullptr;
    b->dirty = true;
    return ADF_OK;
}

template <typename T>
static int batch_build(adf_batch* b) {
    batch_drop_tables(b);
    const int n = (int)b->members.size();
    if (!n) { b->dirty = false; return ADF_OK; }
    std::vector<Fused<T>> fs(2 * n);
    std::vector<TxDev<T>> td(n);
    std::vector<int> it0(n), cur(n);
    std::map<std::tuple<int, int, int, int, int, int, int>, std::array<std::vector<uint2>, TILE_LISTS>> glists;
    // members left on the automatic store hint (ADF_OPT_STORE_NT = -1)
    // store non-temporally when the batch's field buffers exceed the GPU's
    // last-level cache, where ordinary stores would only evict the next
    // step's reads
    int64_t field_bytes = 0;
    for (adf_solver* o : b->members) field_bytes += 8 * o->n3 * (int64_t)o->es;
    const bool auto_nt = field_bytes > LLC_BYTES;
    // total work decides the z-block: enough tiles for every CU
    int64_t cells = 0;
    for (adf_solver* s : b->members) cells += s->n3;
    for (int m = 0; m < n; ++m) {
        adf_solver* s = b->members[m];
        const int ty = sweep_ty_int((int)s->es, fused_cpl(s));
        // auto: 8-plane interior tiles and 4-plane general tiles, halved
        // while the whole batch has fewer than 8 tiles per CU
        int kb = s->block_z;
        if (!kb) {
            kb = 8;
            while (kb > 2) {
                int64_t tiles = 0;
                for (adf_solver* o : b->members)
                    tiles += ((o->dom.ny + ty - 1) / ty) * ((o->dom.nz + kb - 1) / kb) *
                             fused_nseg(o->dom.nx, std::max(fused_cpl(o), 1));
                if (tiles >= 8LL * b->ctx->num_cu) break;
                kb /= 2;
            }
        }
        const int kbg = s->block_z_gen ? s->block_z_gen : 4;
        for (int par = 0; par < 2; ++par) {
            const int save = s->cur;
            s->cur = par;
            fs[2 * m + par] = make_fused<T>(s, ty);
            s->cur = save;
        }
        TxDev<T>& D = td[m];
        D.X = make_txs<T>(s);
        D.fb[0] = (T*)s->fieldb[0];
        D.fb[1] = (T*)s->fieldb[1];
        D.nx = (int)s->dom.nx; D.ny = (int)s->dom.ny; D.nz = (int)s->dom.nz; D.n3 = (int)s->n3;
        D.fuse_mirror = s->fuse_mirror ? 1 : 0;
        it0[m] = (int)(s->next_it - 1);
        cur[m] = s->cur;
        int rc0 = check_tiling(s, ty);
        if (rc0) return rc0;
        std::vector<uint8_t> open;
        if ((rc0 = open_rows(s, open))) return rc0;
        std::vector<uint2> lists[TILE_LISTS];
        split_tiles(s, ty, kb, kbg, 1, (uint32_t)m, open, lists);
        const int cpl = fused_cpl(s);
        const std::tuple<int, int, int, int, int, int, int> key(cpl, ty, s->waves ? s->waves : 1, s->pipeline,
                                                                s->launch == ADF_SWEEP_ONE_LAUNCH ? 1 : 0,
                                                                s->dom.nx % cpl != 0 ? 1 : 0,
                                                                s->pipeline && (s->store_nt > 0 ||
                                                                                (s->store_nt < 0 && auto_nt)) ? 1 : 0);
        auto& G = glists[key];
        for (int c : {LIST_INT, LIST_SIDE}) G[c].insert(G[c].end(), lists[c].begin(), lists[c].end());
    }
    int rc;
    if ((rc = dmalloc(&b->fs, fs.size() * sizeof(Fused<T>), "batch solves"))) return rc;
    if ((rc = dmalloc(&b->td, td.size() * sizeof(TxDev<T>), "batch transceivers"))) return rc;
    if ((rc = dmalloc((void**)&b->it0, n * 4, "batch steps"))) return rc;
    if ((rc = dmalloc((void**)&b->cur, n * 4, "batch parity"))) return rc;
    HIPCHK(hipMemcpy(b->fs, fs.data(), fs.size() * sizeof(Fused<T>), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->td, td.data(), td.size() * sizeof(TxDev<T>), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->it0, it0.data(), n * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->cur, cur.data(), n * 4, hipMemcpyHostToDevice));
    for (auto& kv : glists) {
        adf_batch::Group g{};
        g.cpl = std::get<0>(kv.first);
        g.ty = std::get<1>(kv.first);
        g.waves = b->persistent ? 1 : std::get<2>(kv.first);
        g.pf = std::get<3>(kv.first);
        g.merged = std::get<4>(kv.first);
        g.rag = std::get<5>(kv.first);
        g.store_nt = std::get<6>(kv.first);
        if (g.merged) merge_lists(kv.second[LIST_SIDE], kv.second[LIST_INT], g.waves, kv.second[LIST_ALL]);
        for (int h = 0; h < TILE_LISTS; ++h) {
            const std::vector<uint2>& L = kv.second[h];
            g.n[h] = (int)L.size();
            g.grid[h] = b->persistent ? std::min(g.n[h], 16 * b->ctx->num_cu) : (g.n[h] + g.waves - 1) / g.waves;
            g.tiles[h] = nullptr;
            g.ctr[h] = nullptr;
            if (L.empty()) continue;
            if ((rc = dmalloc((void**)&g.tiles[h], L.size() * sizeof(uint2), "batch tiles"))) return rc;
            if ((rc = dmalloc((void**)&g.ctr[h], 9 * sizeof(unsigned), "batch counters"))) return rc;
            HIPCHK(hipMemcpy(g.tiles[h], L.data(), L.size() * sizeof(uint2), hipMemcpyHostToDevice));
            HIPCHK(hipMemset(g.ctr[h], 0, 9 * sizeof(unsigned)));
        }
        b->groups.push_back(g);
    }
    // transceiver levels of all members, merged level by level
    std::vector<std::vector<uint32_t>> lp, lv;
    std::vector<uint2> fix;
    std::vector<int> mir;
    for (int m = 0; m < n; ++m) {
        adf_solver* s = b->members[m];
        if (!s->fuse_mirror) mir.push_back(m);
        if (!s->ntx) continue;
        const std::vector<uint32_t>& op = s->h_order_p;
        const std::vector<uint32_t>& ov = s->h_order_v;
        if ((int64_t)op.size() != s->ntx || (int64_t)ov.size() != s->ntx)
            return fail(ADF_ESTATE, "adf_batch_run: a member's transceiver schedule is missing");
        auto merge = [&](const std::vector<uint32_t>& o, const std::vector<int>& lev, std::vector<std::vector<uint32_t>>& out) {
            if (out.size() + 1 < lev.size()) out.resize(lev.size() - 1);
            for (size_t L = 0; L + 1 < lev.size(); ++L)
                for (int q = lev[L]; q < lev[L + 1]; ++q) {
                    const uint32_t e = o[q];
                    out[L].push_back(((uint32_t)m << 20) | ((e >> 31) << 19) | (e & 0x7ffffu));
                }
        };
        merge(op, s->lev_p, lp);
        merge(ov, s->lev_v, lv);
        for (uint32_t c : s->h_fix) fix.push_back(make_uint2((uint32_t)m, c));
    }
    auto upload_levels = [&](const std::vector<std::vector<uint32_t>>& lv_, uint32_t** dst, std::vector<int>& off) {
        std::vector<uint32_t> flat;
        off.assign(1, 0);
        for (auto& l : lv_) {
            flat.insert(flat.end(), l.begin(), l.end());
            off.push_back((int)flat.size());
        }
        if (flat.empty()) { off.clear(); return ADF_OK; }
        int r = dmalloc((void**)dst, flat.size() * 4, "batch transceiver levels");
        if (r) return r;
        HIPCHK(hipMemcpy(*dst, flat.data(), flat.size() * 4, hipMemcpyHostToDevice));
        return ADF_OK;
    };
    if ((rc = upload_levels(lp, &b->txp, b->txp_lev))) return rc;
    if ((rc = upload_levels(lv, &b->txv, b->txv_lev))) return rc;
    b->n_fix = (int)fix.size();
    if (b->n_fix) {
        if ((rc = dmalloc((void**)&b->fix, fix.size() * sizeof(uint2), "batch fixups"))) return rc;
        HIPCHK(hipMemcpy(b->fix, fix.data(), fix.size() * sizeof(uint2), hipMemcpyHostToDevice));
    }
    b->n_mirror = (int)mir.size();
    if (b->n_mirror) {
        if ((rc = dmalloc((void**)&b->mirror, mir.size() * 4, "batch mirrors"))) return rc;
        HIPCHK(hipMemcpy(b->mirror, mir.data(), mir.size() * 4, hipMemcpyHostToDevice));
    }
    // every member's transceivers in one launch when each fits the
    // one-launch step (ADF_OPT_TX_LAUNCH not 0, no GF map)
    bool one = true;
    std::vector<TxChain> ch((size_t)n);
    std::vector<int2> map;
    for (int m = 0; m < n && one; ++m) {
        adf_solver* s = b->members[m];
        ch[m] = TxChain{};
        if (!s->ntx) continue;
        if (!s->txc_ok || !s->tx_launch || s->gf_kind) {
            one = false;
            break;
        }
        ch[m] = tx_chain_of(s);
        const int nblk = ch[m].nchain + (ch[m].n_rec + 7) / 8;
        for (int k = 0; k < nblk; ++k) map.push_back(make_int2(m, k));
    }
    if (one && !map.empty()) {
        if ((rc = dmalloc(&b->chains, ch.size() * sizeof(TxChain), "batch transceiver chains")) ||
            (rc = dmalloc((void**)&b->txmap, map.size() * sizeof(int2), "batch transceiver chains")) ||
            (rc = dmalloc((void**)&b->txdone, sizeof(unsigned), "batch transceiver chains")))
            return rc;
        HIPCHK(hipMemcpy(b->chains, ch.data(), ch.size() * sizeof(TxChain), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(b->txmap, map.data(), map.size() * sizeof(int2), hipMemcpyHostToDevice));
        HIPCHK(hipMemset(b->txdone, 0, sizeof(unsigned)));
        b->n_txmap = (int)map.size();
    }
    b->dirty = false;
    return ADF_OK;
}

// The launches of one batched step, all on stream st; every per-solve input
// is read on the device, so the sequence is the same for every step
// (graph-capturable).
template <typename T, bool EXACT>
static void batch_step(adf_batch* b, hipStream_t st) {
    launch_fused_batch<T, EXACT>(b, st);
    const TxDev<T>* td = (const TxDev<T>*)b->td;
    const int nm = (int)b->members.size();
    if (b->n_txmap) {
        const int adv = b->n_mirror == 0;
        hipLaunchKernelGGL((k_tx_step_batch<T, EXACT>), dim3(b->n_txmap), dim3(256), 0, st, (const Fused<T>*)b->fs,
                           td, (const TxChain*)b->chains, b->txmap, b->it0, b->cur, b->txdone, nm, adv);
        if (adv) return;
        hipLaunchKernelGGL((k_mirror_batch<T>), dim3(64, b->n_mirror), dim3(256), 0, st, td, b->cur, b->mirror);
        hipLaunchKernelGGL(k_advance, dim3((nm + 255) / 256), dim3(256), 0, st, b->it0, b->cur, nm);
        return;
    }
    for (size_t L = 0; L + 1 < b->txp_lev.size(); ++L) {
        const int n = b->txp_lev[L + 1] - b->txp_lev[L];
        if (n > 0)
            hipLaunchKernelGGL((k_tx_p_batch<T>), dim3((8 * n + 255) / 256), dim3(256), 0, st, td, b->it0, b->cur,
                               b->txp + b->txp_lev[L], n);
    }
    if (b->n_fix)
        hipLaunchKernelGGL((k_fixup_batch<T, EXACT>), dim3((b->n_fix + 255) / 256), dim3(256), 0, st,
                           (const Fused<T>*)b->fs, b->cur, b->fix, b->n_fix);
    for (size_t L = 0; L + 1 < b->txv_lev.size(); ++L) {
        const int n = b->txv_lev[L + 1] - b->txv_lev[L];
        if (n > 0)
            hipLaunchKernelGGL((k_tx_v_batch<T>), dim3((32 * n + 255) / 256), dim3(256), 0, st, td, b->it0, b->cur,
                               b->txv + b->txv_lev[L], n);
    }
    if (b->n_mirror)
        hipLaunchKernelGGL((k_mirror_batch<T>), dim3(64, b->n_mirror), dim3(256), 0, st, td, b->cur, b->mirror);
    const int n = (int)b->members.size();
    hipLaunchKernelGGL(k_advance, dim3((n + 255) / 256), dim3(256), 0, st, b->it0, b->cur, n);
}

template <typename T, bool EXACT>
static int batch_run_impl(adf_batch* b, int64_t nsteps) {
    hipStream_t st = b->ctx->stream;
    if (b->dirty) {
        int rc = batch_build<T>(b);
        if (rc) return rc;
    } else {
        // members may have been stepped or reset alone: refresh step and parity
        const int n = (int)b->members.size();
        std::vector<int> it0(n), cur(n);
        for (int m = 0; m < n; ++m) { it0[m] = (int)(b->members[m]->next_it - 1); cur[m] = b->members[m]->cur; }
        HIPCHK(hipMemcpyAsync(b->it0, it0.data(), n * 4, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(b->cur, cur.data(), n * 4, hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
    }
    if (!b->ev0) {
        HIPCHK(hipEventCreate(&b->ev0));
        HIPCHK(hipEventCreate(&b->ev1));
    }
    int64_t left = nsteps;
    const int G = b->graph_steps;
    if (G > 0 && left >= G && !b->gexec) {
        HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < G; ++i) batch_step<T, EXACT>(b, st);
        HIPCHK(hipStreamEndCapture(st, &b->graph));
        HIPCHK(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
        b->gexec_steps = G;
    }
    HIPCHK(hipEventRecord(b->ev0, st));
    while (b->gexec && left >= b->gexec_steps) {
        HIPCHK(hipGraphLaunch(b->gexec, st));
        left -= b->gexec_steps;
    }
    for (; left > 0; --left) batch_step<T, EXACT>(b, st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(b->ev1, st));
    HIPCHK(wait_event(b->ev1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->last_step_ms = nsteps > 0 ? ms / (double)nsteps : 0.0;
    for (adf_solver* s : b->members) {
        s->next_it += nsteps;
        if (nsteps & 1) s->cur ^= 1;
    }
    return ADF_OK;
}

int batch_run(adf_batch* b, int64_t nsteps) {
    if (nsteps < 0) return fail(ADF_EINVAL, "adf_batch_run: nsteps < 0");
    if (b->members.empty() || nsteps == 0) return ADF_OK;
    for (adf_solver* s : b->members)
        if (s->next_it + nsteps - 1 > s->dom.nt) return fail(ADF_EINVAL, "adf_batch_run: a member would step past its nt");
    if (b->dtype == ADF_F64) return batch_run_impl<double, true>(b, nsteps);
    if (b->dtype == ADF_F32_REF) return batch_run_impl<float, true>(b, nsteps);
    return batch_run_impl<float, false>(b, nsteps);
}

}  // namespace adf
