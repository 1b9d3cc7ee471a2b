// Served form (include/adfdtd.h adf_server_*): a request loop on its own
// host thread and HIP context.  Requests are admitted in arrival order while
// the device-memory budget allows, become members of the batch of their
// dtype (fields, CPML memory, mask and traces stay resident), and every live
// solve advances together, in chunks of steps, through adf_batch_run; a
// solve that reaches its last step is retired: its traces are copied to
// host memory and its device memory is released, which admits the next
// requests.  Latency is measured from submission to retirement.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/adfdtd.h"
#include "adf_internal.hpp"

using namespace adf;

namespace {

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Request {
    int64_t id = 0;
    // inputs (host copies owned by the server)
    adf_domain dom{};
    int dtype = ADF_F32;
    int64_t npml = 0;
    std::vector<double> coef[8];
    int mask_shape = 0;
    double mask_params[4] = {0, 0, 0, 0};
    std::vector<double> xc, yc, zc;
    std::vector<uint8_t> mask4;
    int64_t ntx = 0;
    std::vector<int64_t> idx;
    std::vector<char> w, src, dir, vn_src, field;
    int64_t start = 1, end = 0;
    uint32_t want = 0;
    int keep_field = 0;
    // state
    int status = ADF_REQ_QUEUED;
    adf_solver* s = nullptr;
    uint64_t bytes = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    std::vector<char> out[10];
    std::vector<char> out_field;
    std::string err;
};

size_t esize(int dtype) { return dtype == ADF_F64 ? 8 : 4; }

// Device bytes a request holds while live: two field buffers, two CPML
// buffers, the mask bit-planes, the ten trace buffers and the packed
// transceiver tables.
uint64_t footprint(const Request& r) {
    const uint64_t nx = r.dom.nx, ny = r.dom.ny, nz = r.dom.nz, n3 = nx * ny * nz, p = r.npml, q = p + 1;
    const uint64_t es = esize(r.dtype);
    const uint64_t npsi = 2 * (p * ny * nz + nx * p * nz + nx * ny * p + q * ny * nz + nx * q * nz + nx * ny * q);
    const uint64_t nwx = (nx + 31) / 32;
    const uint64_t tr = (uint64_t)r.ntx * (uint64_t)r.dom.nt * es;
    return 8 * n3 * es + 2 * npsi * es + 16 * nwx * ny * nz + 10 * tr + (uint64_t)r.ntx * (4 * 3 * 4 + 4 * 8 * es + 16 * es) +
           (1u << 20);
}

}  // namespace

struct adf_server {
    adf_ctx* ctx = nullptr;
    uint64_t budget = 0, in_use = 0, peak = 0;
    int chunk = 64;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> queue;
    std::vector<Request*> live;
    std::map<int64_t, Request*> all;
    std::map<int, adf_batch*> batches;  // one per dtype
    std::thread th;
    bool stop = false;
    int64_t next_id = 1;
    // statistics
    std::vector<double> lat_ms;
    double t_first = -1, t_last = 0, busy_s = 0, lups = 0;
    int64_t completed = 0, failed = 0;
};

namespace {

// Device set-up of an admitted request (server thread).
int setup(adf_server* S, Request* r) {
    adf_cpml cp{};
    cp.npml = r->npml;
    const double** slots[8] = {&cp.a_l, &cp.a_r, &cp.a_hl, &cp.a_hr, &cp.b_l, &cp.b_r, &cp.b_hl, &cp.b_hr};
    for (int c = 0; c < 8; ++c) *slots[c] = r->coef[c].empty() ? nullptr : r->coef[c].data();
    int rc = adf_solver_create(S->ctx, &r->dom, r->dtype, &cp, &r->s);
    if (rc) return rc;
    if (!r->field.empty() && (rc = adf_solver_set_field(r->s, r->field.data()))) return rc;
    if (r->mask_shape == -1) {
        if ((rc = adf_solver_set_mask(r->s, r->mask4.data()))) return rc;
    } else if (r->mask_shape > 0) {
        if ((rc = adf_solver_build_mask(r->s, r->mask_shape, r->mask_params, r->xc.data(), r->yc.data(), r->zc.data())))
            return rc;
    }
    if (r->ntx &&
        (rc = adf_solver_set_transceivers(r->s, r->ntx, r->idx.data(), r->w.data(), r->src.data(), r->dir.data(),
                                          r->vn_src.empty() ? nullptr : r->vn_src.data())))
        return rc;
    r->s->next_it = r->start;
    adf_batch*& b = S->batches[r->dtype];
    if (!b && (rc = adf_batch_create(S->ctx, &b))) return rc;
    return adf_batch_add(b, r->s);
}

// Copy the requested results to host memory and release the device state.
int retire(adf_server* S, Request* r) {
    const size_t tr = (size_t)r->ntx * r->dom.nt * esize(r->dtype);
    int rc = ADF_OK;
    for (int w = 0; w < 10 && !rc; ++w)
        if (((r->want >> w) & 1u) && r->ntx) {
            r->out[w].resize(tr);
            rc = adf_solver_get_trace(r->s, w, r->out[w].data());
        }
    if (!rc && r->keep_field) {
        r->out_field.resize((size_t)4 * r->dom.nx * r->dom.ny * r->dom.nz * esize(r->dtype));
        rc = adf_solver_get_field(r->s, r->out_field.data());
    }
    if (rc) r->err = adf_last_error();
    adf_solver_destroy(r->s);  // leaves its batch
    r->s = nullptr;
    return rc;
}

void loop(adf_server* S) {
    hipSetDevice(S->ctx->device);
    for (;;) {
        std::vector<Request*> admitted;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_work.wait(lk, [&] { return S->stop || !S->queue.empty() || !S->live.empty(); });
            if (S->stop && S->live.empty() && S->queue.empty()) break;
            if (S->live.empty() && !S->stop) {
                // idle server: give a burst of submissions a moment to arrive
                // so that it starts as one batch
                // (until 1 ms passes without a new one, at most 20 ms)
                using clk = std::chrono::steady_clock;
                const auto t_end = clk::now() + std::chrono::milliseconds(20);
                for (;;) {
                    const size_t seen = S->queue.size();
                    const auto until = std::min(clk::now() + std::chrono::milliseconds(1), t_end);
                    if (!S->cv_work.wait_until(lk, until, [&] { return S->stop || S->queue.size() > seen; })) break;
                    if (S->stop || clk::now() >= t_end) break;
                }
            }
            if (S->stop) {  // drop what has not started
                for (Request* r : S->queue) {
                    r->status = ADF_REQ_FAILED;
                    r->err = "server stopped before the request was admitted";
                }
                S->queue.clear();
            }
            // admission in arrival order while the budget allows (a request
            // larger than the whole budget runs alone)
            while (!S->queue.empty()) {
                Request* r = S->queue.front();
                if (S->in_use + r->bytes > S->budget && !(S->live.empty() && admitted.empty())) break;
                S->queue.pop_front();
                S->in_use += r->bytes;
                S->peak = std::max(S->peak, S->in_use);
                r->status = ADF_REQ_RUNNING;
                r->t_admit = now_s();
                admitted.push_back(r);
            }
        }
        for (Request* r : admitted) {
            const int rc = setup(S, r);
            std::lock_guard<std::mutex> lk(S->mu);
            if (rc) {
                r->err = adf_last_error();
                if (r->s) { adf_solver_destroy(r->s); r->s = nullptr; }
                r->status = ADF_REQ_FAILED;
                r->t_done = now_s();
                S->in_use -= r->bytes;
                S->failed++;
                S->cv_done.notify_all();
            } else {
                S->live.push_back(r);
            }
        }
        if (S->live.empty()) continue;
        // one chunk: every live solve advances by the same number of steps,
        // at most up to the nearest last step
        int64_t n = S->chunk;
        for (Request* r : S->live) n = std::min(n, r->end - r->s->next_it + 1);
        const double t0 = now_s();
        double lups = 0;
        int rc = ADF_OK;
        for (auto& kv : S->batches) {
            int64_t m = 0;
            adf_batch_size(kv.second, &m);
            if (!m) continue;
            if ((rc = adf_batch_run(kv.second, n))) break;
        }
        for (Request* r : S->live) lups += (double)r->dom.nx * r->dom.ny * r->dom.nz * n;
        const double t1 = now_s();
        std::vector<Request*> done, keep;
        for (Request* r : S->live) (rc || r->s->next_it > r->end ? done : keep).push_back(r);
        const std::string msg = rc ? std::string(adf_last_error()) : std::string();
        std::vector<int> rrc(done.size());
        for (size_t i = 0; i < done.size(); ++i) rrc[i] = rc ? rc : retire(S, done[i]);
        uint64_t headroom = 0;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->busy_s += t1 - t0;
            S->lups += lups;
            S->live = keep;
            const double t = now_s();
            for (size_t i = 0; i < done.size(); ++i) {
                Request* r = done[i];
                r->t_done = t;
                S->in_use -= r->bytes;
                if (rrc[i]) {
                    if (r->s) { adf_solver_destroy(r->s); r->s = nullptr; }
                    if (r->err.empty()) r->err = msg;
                    r->status = ADF_REQ_FAILED;
                    S->failed++;
                } else {
                    r->status = ADF_REQ_DONE;
                    S->completed++;
                    S->lat_ms.push_back((r->t_done - r->t_submit) * 1e3);
                    S->t_last = t;
                }
            }
            S->cv_done.notify_all();
            headroom = S->budget > S->in_use ? S->budget - S->in_use : 0;
        }
        // the blocks the retired requests freed stay pooled for the next
        // admissions only within the budget's headroom
        if (!done.empty()) device_pool_trim(S->ctx->device, (size_t)headroom);
    }
}

template <typename V>
void copy_in(V& dst, const void* p, size_t bytes) {
    dst.resize(bytes / sizeof(typename V::value_type));
    if (bytes) std::memcpy(dst.data(), p, bytes);
}

}  // namespace

extern "C" int adf_server_create(int device, uint64_t hbm_budget_bytes, adf_server** out) {
    if (!out) return fail(ADF_EINVAL, "adf_server_create: null out");
    adf_ctx* ctx = nullptr;
    int rc = adf_ctx_create(device, &ctx);
    if (rc) return rc;
    adf_server* S = new adf_server();
    S->ctx = ctx;
    S->budget = hbm_budget_bytes;
    // blocks pooled by earlier solves are released: the budget bounds live
    // requests plus what the pool keeps for them (see the loop)
    device_pool_release(device);
    if (!S->budget) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) == hipSuccess) S->budget = (uint64_t)(0.85 * (double)fr);
    }
    S->th = std::thread(loop, S);
    *out = S;
    return ADF_OK;
}

extern "C" int adf_server_destroy(adf_server* S) {
    if (!S) return ADF_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    if (S->th.joinable()) S->th.join();
    for (auto& kv : S->batches) adf_batch_destroy(kv.second);
    for (auto& kv : S->all) {
        if (kv.second->s) adf_solver_destroy(kv.second->s);
        delete kv.second;
    }
    adf_ctx_destroy(S->ctx);
    delete S;
    return ADF_OK;
}

extern "C" int adf_server_set_option(adf_server* S, int option, int64_t value) {
    if (!S) return fail(ADF_EINVAL, "adf_server_set_option: null server");
    std::lock_guard<std::mutex> lk(S->mu);
    if (option == ADF_SERVER_CHUNK) {
        if (value < 1 || value > 100000) return fail(ADF_EINVAL, "ADF_SERVER_CHUNK: 1..100000");
        S->chunk = (int)value;
    } else if (option == ADF_SERVER_BUDGET) {
        if (value < 0) return fail(ADF_EINVAL, "ADF_SERVER_BUDGET: >= 0");
        S->budget = (uint64_t)value;
    } else {
        return fail(ADF_EINVAL, "adf_server_set_option: unknown option");
    }
    return ADF_OK;
}

extern "C" int adf_server_submit(adf_server* S, const adf_request* q, int64_t* id) {
    if (!S || !q || !id) return fail(ADF_EINVAL, "adf_server_submit: null argument");
    if (q->dtype != ADF_F32 && q->dtype != ADF_F64 && q->dtype != ADF_F32_REF)
        return fail(ADF_EINVAL, "adf_server_submit: bad dtype");
    const adf_domain& d = q->dom;
    if (d.nx < 3 || d.ny < 3 || d.nz < 3 || d.nt < 1) return fail(ADF_EINVAL, "adf_server_submit: bad domain");
    // the batched (fused) sweep's limits: one buffer descriptor per field
    // buffer, at most 63 row segments of 62 x 4 cells
    const int64_t wb = q->dtype == ADF_F64 ? 8 : 4;
    if (4 * d.nx * d.ny * d.nz * wb >= (int64_t(1) << 32) || d.nx > 63 * 62 * 4)
        return fail(ADF_EINVAL, "adf_server_submit: served solves need 4 nx ny nz sizeof(T) < 4 GiB and nx <= 15624");
    if (q->start_step < 1 || q->end_step > d.nt || q->start_step > q->end_step)
        return fail(ADF_EINVAL, "adf_server_submit: steps must satisfy 1 <= start <= end <= nt");
    if (q->mask_shape < -1 || q->mask_shape > 4) return fail(ADF_EINVAL, "adf_server_submit: bad mask_shape");
    if (q->mask_shape == -1 && !q->mask4) return fail(ADF_EINVAL, "adf_server_submit: mask4 missing");
    if (q->mask_shape > 0 && (!q->xc || !q->yc || !q->zc)) return fail(ADF_EINVAL, "adf_server_submit: coordinates missing");
    if (q->ntx < 0 || (q->ntx > 0 && (!q->idx || !q->w || !q->src || !q->dir)))
        return fail(ADF_EINVAL, "adf_server_submit: transceiver arrays missing");
    if (q->cpml.npml < 0 || (q->cpml.npml > 0 && (!q->cpml.a_l || !q->cpml.b_l)))
        return fail(ADF_EINVAL, "adf_server_submit: CPML coefficients missing");
    Request* r = new Request();
    r->dom = d;
    r->dtype = q->dtype;
    r->npml = q->cpml.npml;
    const double* cs[8] = {q->cpml.a_l, q->cpml.a_r, q->cpml.a_hl, q->cpml.a_hr, q->cpml.b_l, q->cpml.b_r, q->cpml.b_hl, q->cpml.b_hr};
    const int64_t len[8] = {r->npml, r->npml, r->npml + 1, r->npml + 1, r->npml, r->npml, r->npml + 1, r->npml + 1};
    if (r->npml > 0)
        for (int c = 0; c < 8; ++c) copy_in(r->coef[c], cs[c], (size_t)len[c] * 8);
    r->mask_shape = q->mask_shape;
    std::memcpy(r->mask_params, q->mask_params, sizeof(r->mask_params));
    if (q->mask_shape > 0) {
        copy_in(r->xc, q->xc, (size_t)d.nx * 8);
        copy_in(r->yc, q->yc, (size_t)d.ny * 8);
        copy_in(r->zc, q->zc, (size_t)d.nz * 8);
    } else if (q->mask_shape == -1) {
        copy_in(r->mask4, q->mask4, (size_t)4 * d.nx * d.ny * d.nz);
    }
    const size_t es = esize(q->dtype);
    r->ntx = q->ntx;
    if (q->ntx) {
        copy_in(r->idx, q->idx, (size_t)q->ntx * 12 * 8);
        copy_in(r->w, q->w, (size_t)q->ntx * 32 * es);
        copy_in(r->src, q->src, (size_t)q->ntx * 4 * d.nt * es);
        copy_in(r->dir, q->dir, (size_t)q->ntx * 4 * es);
        if (q->vn_src) copy_in(r->vn_src, q->vn_src, (size_t)q->ntx * d.nt * es);
    }
    if (q->field) copy_in(r->field, q->field, (size_t)4 * d.nx * d.ny * d.nz * es);
    r->start = q->start_step;
    r->end = q->end_step;
    r->want = q->want ? q->want : 0x1fu;
    r->keep_field = q->keep_field;
    r->bytes = footprint(*r);
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->stop) {
            delete r;
            return fail(ADF_ESTATE, "adf_server_submit: server is stopping");
        }
        r->id = S->next_id++;
        r->t_submit = now_s();
        if (S->t_first < 0) S->t_first = r->t_submit;
        S->all[r->id] = r;
        S->queue.push_back(r);
        *id = r->id;
    }
    S->cv_work.notify_all();
    return ADF_OK;
}

extern "C" int adf_server_poll(adf_server* S, int64_t id, int* status) {
    if (!S || !status) return fail(ADF_EINVAL, "adf_server_poll: null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) return fail(ADF_EINVAL, "adf_server_poll: unknown request id");
    *status = it->second->status;
    return ADF_OK;
}

extern "C" int adf_server_wait(adf_server* S, int64_t id, double timeout_s, int* status) {
    if (!S || !status) return fail(ADF_EINVAL, "adf_server_wait: null argument");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) return fail(ADF_EINVAL, "adf_server_wait: unknown request id");
    Request* r = it->second;
    auto fin = [&] { return r->status == ADF_REQ_DONE || r->status == ADF_REQ_FAILED; };
    if (timeout_s < 0) S->cv_done.wait(lk, fin);
    else S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), fin);
    *status = r->status;
    return ADF_OK;
}

extern "C" int adf_server_get_trace(adf_server* S, int64_t id, int which, void* out) {
    if (!S || !out) return fail(ADF_EINVAL, "adf_server_get_trace: null argument");
    if (which < 0 || which > 9) return fail(ADF_EINVAL, "adf_server_get_trace: bad buffer id");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) return fail(ADF_EINVAL, "adf_server_get_trace: unknown request id");
    Request* r = it->second;
    if (r->status != ADF_REQ_DONE) return fail(ADF_ESTATE, "adf_server_get_trace: request not done");
    if (!((r->want >> which) & 1u)) return fail(ADF_EINVAL, "adf_server_get_trace: buffer not requested");
    if (!r->out[which].empty()) std::memcpy(out, r->out[which].data(), r->out[which].size());
    return ADF_OK;
}

extern "C" int adf_server_get_field(adf_server* S, int64_t id, void* out) {
    if (!S || !out) return fail(ADF_EINVAL, "adf_server_get_field: null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) return fail(ADF_EINVAL, "adf_server_get_field: unknown request id");
    Request* r = it->second;
    if (r->status != ADF_REQ_DONE || r->out_field.empty())
        return fail(ADF_ESTATE, "adf_server_get_field: request not done or submitted without keep_field");
    std::memcpy(out, r->out_field.data(), r->out_field.size());
    return ADF_OK;
}

extern "C" const char* adf_server_request_error(adf_server* S, int64_t id) {
    if (!S) return "";
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    return it == S->all.end() ? "" : it->second->err.c_str();
}

extern "C" int adf_server_release(adf_server* S, int64_t id) {
    if (!S) return fail(ADF_EINVAL, "adf_server_release: null server");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) return fail(ADF_EINVAL, "adf_server_release: unknown request id");
    Request* r = it->second;
    if (r->status != ADF_REQ_DONE && r->status != ADF_REQ_FAILED)
        return fail(ADF_ESTATE, "adf_server_release: request still queued or running");
    delete r;
    S->all.erase(it);
    return ADF_OK;
}

extern "C" int adf_server_stats(adf_server* S, adf_server_info* st) {
    if (!S || !st) return fail(ADF_EINVAL, "adf_server_stats: null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    std::memset(st, 0, sizeof(*st));
    st->completed = S->completed;
    st->failed = S->failed;
    st->running = (int64_t)S->live.size();
    st->queued = (int64_t)S->queue.size();
    st->bytes_in_use = S->in_use;
    st->peak_bytes = S->peak;
    st->budget_bytes = S->budget;
    st->busy_s = S->busy_s;
    st->glups = S->busy_s > 0 ? S->lups / S->busy_s / 1e9 : 0.0;
    std::vector<double> v = S->lat_ms;
    if (!v.empty()) {
        std::sort(v.begin(), v.end());
        auto pct = [&](double f) {  // nearest rank
            size_t k = (size_t)std::ceil(f * (double)v.size());
            return v[std::min(v.size() - 1, k ? k - 1 : 0)];
        };
        st->p50_ms = pct(0.50);
        st->p99_ms = pct(0.99);
        double m = 0;
        for (double x : v) m += x;
        st->mean_ms = m / (double)v.size();
        const double span = S->t_last - S->t_first;
        st->solves_per_s = span > 0 ? (double)v.size() / span : 0.0;
    }
    return ADF_OK;
}
