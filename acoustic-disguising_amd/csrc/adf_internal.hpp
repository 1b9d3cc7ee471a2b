// Internal state of the adfdtd C ABI (not part of the public header).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/adfdtd.h"

namespace adf {
extern thread_local std::string g_err;
int fail(int code, const std::string& msg);
}  // namespace adf

#define HIPCHK(expr)                                                                        \
    do {                                                                                    \
        hipError_t _e = (expr);                                                             \
        if (_e != hipSuccess)                                                               \
            return adf::fail(ADF_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

// Waits for event e by polling it, so that a run returns as soon as its
// last step completes rather than when a sleeping thread is woken.
inline hipError_t wait_event(hipEvent_t e) {
    hipError_t r;
    while ((r = hipEventQuery(e)) == hipErrorNotReady) {
    }
    return r;
}

struct adf_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // second stream of the same (default) priority: with
    // ADF_SWEEP_TWO_STREAMS the side tiles of a sweep run on it, at the same
    // time as the interior tiles on `stream`
    hipStream_t side = nullptr;
    int num_cu = 256;
};

struct adf_batch;

struct adf_solver {
    adf_ctx* ctx = nullptr;
    adf_domain dom{};
    int dtype = ADF_F32;
    size_t es = 4;
    int64_t p = 0;
    int64_t n3 = 0;
    int64_t nwx = 0;

    // field (nx,ny,nz,4) and CPML memory (reference layouts), two buffers
    // each: the fused pass reads buffer `cur` and writes buffer `cur ^ 1`
    void* fieldb[2] = {nullptr, nullptr};
    void* mvb[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
    void* mpb[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
    void* psib[2] = {nullptr, nullptr};  // one allocation per buffer holding mv* then mp*
    int64_t psi_off[6] = {0, 0, 0, 0, 0, 0};  // element offsets of mv_x..mp_z in psib
    int64_t psi_n = 0;
    int cur = 0;
    int64_t mv_n[3] = {0, 0, 0};
    int64_t mp_n[3] = {0, 0, 0};
    // a (p,2), b (p,2), ah (p+1,2), bh (p+1,2) in the storage type
    void* coef = nullptr;
    // 4 mask bit-planes
    uint32_t* mask = nullptr;
    bool has_mask = false;
    // rows (j,k) whose four mask planes are all open, kept from the last
    // scan of the mask; cleared by every mask change
    std::vector<uint8_t> open_h;
    bool open_ok = false;

    // transceivers (device copies; traces in the reference (nt,ntx) layout)
    int64_t ntx = 0;
    int32_t* tx_idx = nullptr;
    void* tx_w = nullptr;
    void* tx_src = nullptr;
    void* tx_rec = nullptr;
    void* tx_dir = nullptr;
    void* tx_vn_src = nullptr;
    void* tx_vn_rec = nullptr;
    uint32_t* order_p = nullptr;  // entries: index | injector << 31
    uint32_t* order_v = nullptr;
    std::vector<uint32_t> h_order_p, h_order_v;  // host copies of order_p, order_v
    std::vector<int> lev_p;  // level offsets into order_p (size nlev+1)
    // host copies for rebuilding the schedule: 1-based int32 idx [4][ntx][3]
    // and per-transceiver injector flags by source buffer
    std::vector<int32_t> h_idx;
    std::vector<uint8_t> src_inj_p, src_inj_v[3], vn_inj, gf_inj;
    std::vector<int> lev_v;
    bool fuse_mirror = true;
    // velocity cells next to pressure-injection stencils (fused path)
    uint32_t* fix_cells = nullptr;
    int n_fix = 0;
    std::vector<uint32_t> h_fix;  // host copy of fix_cells
    // one-launch transceiver step (k_tx_step): the chain blocks' lists
    // followed by the record-only indices (from txc_rec_off), and the chain
    // blocks' offsets (TXC_HDR ints each)
    uint32_t* txc_list = nullptr;
    int* txc_hdr = nullptr;
    int txc_nchain = 0, txc_nrec = 0, txc_rec_off = 0;
    bool txc_ok = false;  // the chain fits k_tx_step's limits
    std::vector<uint32_t> txc_list_h;  // host copies of txc_list / txc_hdr
    std::vector<int> txc_hdr_h;
    // transceivers carried by the sweep launch (ADF_OPT_TX_LAUNCH = 2): the
    // all-tiles list with the chain's tiles marked (TILE_SIGNAL) and moved to
    // the front of their XCD chunks, per chain block the tiles it waits for
    // (txs_wait_off: nchain + 1 offsets into txs_wait), per tile the epoch of
    // its last store, a wait-ran-out flag.  txs_state: 1 built for the
    // current tiles and chain, 0 stale, -1 not applicable.
    std::vector<uint2> tiles_all_h;  // host copy of the all-tiles list
    uint2* txs_tiles = nullptr;
    int* txs_wait = nullptr;
    int* txs_wait_off = nullptr;
    int* txs_flag = nullptr;
    int* txs_err = nullptr;
    int* txs_err_h = nullptr;  // pinned host copy of *txs_err, read after a run
    int txs_state = 0;
    int txs_epoch = 0;
    bool txs_pending = false;  // the last step's records are not taken yet

    // fused-sweep tile lists for tiling (tl_ty, tl_kb, tl_kbg, tl_w):
    // interior tiles, side tiles (CPML-slab and boundary) and all tiles
    // (LIST_* order)
    int tl_ty = 0, tl_kb = 0, tl_kbg = 0, tl_w = 0;
    uint2* tiles[3] = {nullptr, nullptr, nullptr};
    int ntl[3] = {0, 0, 0};
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;

    // kernel path and tiling (adf_solver_set_option)
    int path = 0;      // 0 auto, 1 two-pass, 2 fused
    int tile_y = 0;    // 0 auto
    int block_z = 0;   // 0 auto
    int waves = 0;     // waves per workgroup of the fused kernel (0 auto)
    int block_z_gen = 0;  // planes per general tile (0 auto)
    int pipeline = 1;     // pipelined (prefetching) tile routines
    int store_nt = -1;    // non-temporal field stores in the pipelined routines: 1, 0, -1 automatic
    int launch = 0;       // ADF_SWEEP_* (one launch, two streams, two serial)
    int tx_launch = 2;    // ADF_OPT_TX_LAUNCH
    int cpl_opt = 0;      // ADF_OPT_CELLS_PER_LANE (0 auto)

    // GF extrapolation map
    int gf_kind = 0;
    int gf_wb = 1;
    int64_t gf_ng = 0, gf_d2 = 0, gf_d3 = 0;
    void* gf_t[4] = {nullptr, nullptr, nullptr, nullptr};   // far-pass tile-major copies (gf_tile_index)
    void* gf_tn[4] = {nullptr, nullptr, nullptr, nullptr};  // lags 0..GF_M-1 as [s][r][GF_M]
    int32_t* gf_irec = nullptr;
    int32_t* gf_isrc = nullptr;
    int64_t gf_nr = 0, gf_ns = 0;
    double* gf_h0 = nullptr;
    double* gf_h1 = nullptr;
    // blocked gather: far-pass partial sums [chunk][GF_M][2][n_isrc] of the
    // block of GF_M steps starting at step gf_B (0-based; -1: none)
    double* gf_part = nullptr;
    int gf_rc = 1;
    int gf_nchunk = 0;
    int64_t gf_max_index = 0;          // largest 1-based iRec/iSrc of the map
    std::vector<int64_t> gf_isrc_h;    // the map's iSrc (1-based), host copy
    std::vector<int64_t> gf_irec_h;    // the map's iRec (1-based), host copy
    int32_t* gf_slot_first = nullptr;  // per transceiver: first input slot reading it (-1: none)
    int32_t* gf_slot_next = nullptr;   // per input slot: next slot reading the same transceiver
    int64_t gf_B = -1;
    // the history (gf_h0/gf_h1) equals the one computed from the record
    // traces at times [0, gf_hist_valid)
    int64_t gf_hist_valid = 0;

    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_step_ms = 0.0;
    int64_t next_it = 1;          // next step a batch run takes (1-based)
    adf_batch* batch = nullptr;   // batch this solver belongs to, if any
    // optional per-launch timing of the field-update kernel
    int time_sweep = 0;           // every time_sweep-th step (0: off)
    int64_t sweep_seq = 0;        // steps of the current run so far
    std::vector<hipEvent_t> sweep_ev;  // pairs (start, stop)
    int sweep_used = 0;
    double sweep_avg_ms = 0.0;
    int64_t sweep_n = 0;
};

// A set of solvers advanced together (include/adfdtd.h adf_batch_*).
struct adf_batch {
    adf_ctx* ctx = nullptr;
    std::vector<adf_solver*> members;
    int dtype = 0;
    bool dirty = true;
    bool persistent = false;
    int graph_steps = 16;     // steps per captured graph (0: no graph)
    // device tables (rebuilt when dirty)
    void* fs = nullptr;       // Fused<T>[2 * n]: [2s + cur] = (old = cur, new = cur ^ 1)
    void* td = nullptr;       // TxDev<T>[n]
    int* it0 = nullptr;       // per member: 0-based step index of the next step
    int* cur = nullptr;       // per member: buffer holding the current field
    // one group per (cells per lane, rows per wave, waves per workgroup);
    struct Group {
        int cpl, ty, waves, pf;  // pf: pipelined tile routines
        int rag;                 // rows not a multiple of cpl long
        int store_nt;            // non-temporal field stores (pipelined routines)
        int merged;             // one launch over LIST_ALL, else LIST_SIDE then LIST_INT
        int n[3], grid[3];      // per tile list (LIST_* order)
        uint2* tiles[3];
        unsigned* ctr[3];       // 8 per-XCD queue heads + a finished-wave count
    };
    std::vector<Group> groups;
    uint32_t* txp = nullptr;
    std::vector<int> txp_lev;
    uint32_t* txv = nullptr;
    std::vector<int> txv_lev;
    uint2* fix = nullptr;
    int n_fix = 0;
    int* mirror = nullptr;
    int n_mirror = 0;
    // one launch for every member's transceivers (k_tx_step_batch): each
    // member's chain lists, per block (member, block of its step), and the
    // launch's finished-block count; n_txmap 0: the per-level launches
    void* chains = nullptr;
    int2* txmap = nullptr;
    int n_txmap = 0;
    unsigned* txdone = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int gexec_steps = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_step_ms = 0.0;
};

namespace adf {
// Frees the device allocations pooled for reuse on device dev.
void device_pool_release(int dev);
// Frees pooled allocations of device dev until at most keep bytes stay pooled.
void device_pool_trim(int dev, size_t keep);
// A change to a member's schedule, mask or tiling invalidates its batch's
// device tables.
inline void touch(adf_solver* s) {
    s->tl_ty = 0;  // tile lists carry mask-derived flags: rebuild
    if (s->batch) s->batch->dirty = true;
}
int batch_add(adf_batch* b, adf_solver* s);
int batch_remove(adf_batch* b, adf_solver* s);
int batch_run(adf_batch* b, int64_t nsteps);
void batch_free(adf_batch* b);
int solver_alloc(adf_solver* s, const adf_cpml* cp);
void solver_free(adf_solver* s);
int solver_zero_state(adf_solver* s);
int solver_set_txs(adf_solver* s, int64_t ntx, const int64_t* idx, const void* w, const void* src,
                   const void* dir, const void* vn_src);
void* solver_trace_ptr(adf_solver* s, int which);
int solver_set_gf(adf_solver* s, int kind, const double* const* tensors, int64_t nt_gf, int64_t dim2,
                  int64_t dim3, const int64_t* iRec, int64_t n_irec, const int64_t* iSrc, int64_t n_isrc,
                  int writeback);
int solver_run(adf_solver* s, int64_t start_step, int64_t end_step);
int resample_scratch_init(int device);
int solver_slice(adf_solver* s, int axis, int64_t index, void* out);
void* cur_field(adf_solver* s);
void* cur_mv(adf_solver* s, int a);
void* cur_mp(adf_solver* s, int a);
int solver_set_option(adf_solver* s, int option, int64_t value);
int solver_sweep_stats(adf_solver* s, double* avg_ms, int64_t* n);
int solver_trace_written(adf_solver* s, int which, const void* data);
bool fused_supported(const adf_solver* s);
void pick_tiles(const adf_solver* s, int* ty, int* kb, int* kbg);
int solver_build_mask_device(adf_solver* s, int shape, const double* params, const double* xc,
                             const double* yc, const double* zc);
}  // namespace adf
