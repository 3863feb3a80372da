// Host-side objects of the placer runtime.
#pragma once
#include <hip/hip_runtime.h>

#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "ep_internal.h"

struct ep_ctx {};

namespace ep {

constexpr int kStatePool = 8192;      // resident designs per context
// Transform engines: 0 MFMA with LDS staging in K stages of 32, 1 FP32/FP64
// FMA, 2 MFMA split-K inside a block, 3 MFMA with LDS staging in K stages of
// 64, 4 engine 0 with products of K >= 512 split across blocks into K slices
// of 128 (partial planes summed by the next pass).  Engine 5 (kEngineFft)
// runs every pass as batched row FFTs (FAST path, power-of-two bin counts).
constexpr int kEngines = 5;      // dense-product engines (job lists per engine)
constexpr int kEngineFft = 5;
constexpr int kNumEngines = 6;
constexpr int kDefaultLogCap = 2048;  // iteration records per design

// Constant transform matrices of one size and field mode: [0] C2, [1] C3,
// [2] S3 (or the IDXST sine matrix), each in FP64, FP32 and BF16 hi/lo.
struct Transform {
    int n = 0, mode = 0;
    void* mem = nullptr;
    double* d[3] = {nullptr, nullptr, nullptr};
    float* f[3] = {nullptr, nullptr, nullptr};
    void* hi[3] = {nullptr, nullptr, nullptr};
    void* lo[3] = {nullptr, nullptr, nullptr};
    float* fft_tw = nullptr;        // float2[n]: twiddles of the row passes' split (fft_twiddle_table)
    float* fft_tw_solve = nullptr;  // float2[n]: twiddles of the column solve's split
    float* dct_tw = nullptr;  // float2[n]: exp(-i pi k / 2n)
};

struct Design {
    uint32_t id = 0;
    int prec = EP_PREC_FAST;
    int log_cap = kDefaultLogCap;
    int state_idx = -1;
    void* arena = nullptr;
    size_t bytes = 0;
    size_t hf_pin_count = 0;
    size_t cs_count = 0;            // movable pins of nets of <= 64 pins (cell -> slot CSR entries)
    size_t chf_count = 0;           // movable pins of high-fanout nets
    bool weighted = false;          // per-net weights given (slot_w allocated)
    std::vector<double> fixed0[2];  // the fixed cells' uploaded locations (restored after a
                                    // best-snapshot placement moved cluster children)
    std::vector<uint32_t> net_off_h, pin_cell_h;  // host copy of the net CSR (ep_design_upload
                                                  // only: the legalized HPWL reads it)
    DesignDev dev{};
};

// A batch of resident designs sharing every launch of an iteration.  The
// transform jobs of each pass are split by engine ([pass][engine]); a
// design's engine depends only on its own shape (and the context's tuning
// table), so it takes the same path in any batch.
struct Batch {
    std::vector<Design*> ds;
    std::vector<ep_config> cfgs;
    int prec = EP_PREC_FAST;
    int log_hpwl = 1;
    int cell_blocks = 0, bin_blocks = 0, wa_blocks = 0, hf_blocks = 0, scat_blocks = 0;
    int tile_blocks = 0;
    int njobs2[4][kEngines] = {};
    int tiles2[4][kEngines] = {};
    GemmJob* jobs2[4][kEngines] = {};
    int32_t* bjob2[4][kEngines] = {};  // job of each block of a pass (one load at block start)
    int nfft[4] = {}, fft_blocks[4] = {};  // FFT engine: jobs and blocks of each pass
    FftJob* fjobs[4] = {};
    int32_t* fblk[4] = {};                 // (job, first row) of each block
    PlaneSum* sums = nullptr;        // split final passes (device), summed after pass 3
    int n_sums = 0, sum_blocks = 0;
    void* mem = nullptr;        // descriptors, prefixes, jobs
    bool ext_mem = false;       // mem is a caller-owned device buffer (not freed here)
    DesignDev* d_dev = nullptr;
    const int32_t* dzero = nullptr;  // a device zero: prefix of a one-design view
    BatchDev bd{};
    hipGraphExec_t graph[2] = {nullptr, nullptr};  // replayed chunk without / with the fix pass
    int graph_iters[2] = {0, 0};
    // keep_graphs: a rebuild keeps the instantiated graphs, marked stale; the
    // next launch captures the new chunk and updates them in place
    // (hipGraphExecUpdate), instantiating only when the launches changed
    bool keep_graphs = false;
    bool graph_stale[2] = {false, false};
    std::vector<double> done_ms;     // per design: device time from its reset to its stop
};

struct Server;
struct ServerDeleter {
    void operator()(Server* s) const;
};

struct Ctx : ep_ctx {
    int device = 0;
    uint64_t budget = 0;
    uint64_t used = 0;
    uint64_t transform_bytes = 0;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;      // WA gradient, concurrent with the density chain
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    hipEvent_t ev_states = nullptr;  // the states' copy behind a queued chunk has landed
    double wall_khz = 0.0;           // device wall-clock rate (wall_clock64 ticks per ms)
    std::recursive_mutex mu;     // serialises device work
    std::mutex err_mu;
    std::string last_error, err_copy;
    std::map<uint32_t, std::unique_ptr<Design>> designs;
    std::map<std::pair<int, int>, std::unique_ptr<Transform>> transforms;
    DevState* state_pool = nullptr;
    DevState* state_host = nullptr;  // pinned mirror
    std::vector<int> free_states;
    uint32_t next_id = 1;
    double t0 = 0.0;
    std::unique_ptr<Batch> cache;    // last built batch (graph kept for replays)
    std::multimap<size_t, void*> arena_pool;  // freed arenas by size, reusable
    uint64_t pooled = 0;                       // bytes held by arena_pool
    ep_config cache_cfg{};
    bool cache_cfg_valid = false;
    ep_config eval_cfg{};            // configuration of the ep_eval_* calls
    bool server_running = false;
    int graph_captures = 0;          // hipGraphs instantiated by batch_run_iters
    int graph_updates = 0;           // instantiated hipGraphs updated in place instead
    std::map<std::pair<int, int>, int> tune;   // (Bx, By) -> transform engine (autotuned)
    std::unique_ptr<Server, ServerDeleter> server;
};

int fail(Ctx* c, int code, const std::string& msg);
double now_ms();
int get_transform(Ctx* c, int n, int mode, Transform** out);
// Net slots of a design (nets of <= 64 pins in groups of 2..64, padded to
// whole waves), its high-fanout nets and the cell -> slot CSR.
struct NetLayout {
    std::vector<uint32_t> slots;
    int n_buckets = kMaxBuckets;
    int bucket_wave_start[kMaxBuckets + 1] = {0};
    int bucket_group[kMaxBuckets] = {0};
    std::vector<uint32_t> hf_off, hf_pins;
    std::vector<uint32_t> cs_off, cs_idx;
    std::vector<uint32_t> cs8;      // [cell][kCellSlots]: first slots of each cell, then the zero slot
    std::vector<int32_t> slot_net;  // net of each slot of `slots` (-1: padding)
    std::vector<int32_t> hf_net;    // net of each high-fanout net
    std::vector<uint32_t> chf_off, chf_net;  // cell -> high-fanout nets it has pins on
};

// Everything an upload computes on the host (layout, sizes, the initial
// contents of the arena in the arithmetic type), so it can run on the
// submitting thread while the device keeps working.  The vectors must stay
// alive until the stream that copies them has been synchronised.
struct Prepared {
    std::unique_ptr<Design> D;  // sizes filled in, no arena yet
    size_t bytes = 0;           // arena size
    NetLayout L;
    std::vector<char> x0[2], fixed[2], pinw, spec[2];  // arithmetic type T
    std::vector<char> fixed_xy;  // fixed locations, x and y interleaved (tail of look)
    std::vector<char> slot_w;   // T per slot: its net's weight (empty: all nets weigh 1)
    std::vector<float> hs;
    std::vector<uint32_t> ident;
    std::vector<uint32_t> order;  // the cells by the tile of their initial position
    std::vector<float> order_hs;  // their half sides, in that order
};

// Validates d and fills P (host only; c receives the error, may be null).
int prepare_design(Ctx* c, const ep_design* d, int prec, int log_cap, Prepared& P);
// Allocates the arena (from the pool when it can), queues the zeroing and
// the copies of P on the context stream and registers the design.  Caller
// holds c->mu and keeps P alive until the stream is synchronised.
int commit_design(Ctx* c, Prepared& P, uint32_t* id_out);
int upload_locked(Ctx* c, const ep_design* d, int prec, int log_cap, uint32_t* id_out);
// Device arenas are pooled: a freed arena is kept (up to a cap) for the next
// design of about its size, so serving does not pay hipMalloc/hipFree (and
// the device synchronisation of hipFree) per request.
int arena_alloc(Ctx* c, size_t bytes, void** out, size_t* got);
void arena_release(Ctx* c, void* p, size_t bytes);
void arena_pool_clear(Ctx* c);
int free_design_locked(Ctx* c, uint32_t id);
void release_batch_cache(Ctx* c);

// Batches (ep_place.cpp)
int batch_build(Ctx* c, const std::vector<Design*>& ds, const ep_config* cfg, Batch& b);
int batch_build_cfgs(Ctx* c, const std::vector<Design*>& ds, const std::vector<const ep_config*>& cfgs,
                     Batch& b, void* ext_mem, size_t ext_cap);
int batch_reset_some(Ctx* c, Batch& b, const std::vector<int>& idx);
int validate_config(Ctx* c, const ep_config* cfg, int log_cap);
// Runs a built batch until every design stops (graph-replayed chunks, or
// eager launches without use_graph).
int batch_place(Ctx* c, Batch& b, int sync_every, bool use_graph);
// The batch cached for one design list and configuration (rebuilt on a change).
Batch* cached_batch(Ctx* c, const std::vector<Design*>& ds, const ep_config* cfg, int* rc);
// Writes the fixed cells' locations of D (both copies the kernels read).
int write_fixed(Ctx* c, Design& D, const double* fx, const double* fy);
Design* find_design(Ctx* c, uint32_t id);
// Device address of the allocation-free entry point's buffer on the
// context's device (ep_noalloc.hip); called once at context creation.
int noalloc_init(Ctx* c);
void batch_release(Batch& b);
int batch_reset(Ctx* c, Batch& b, const ep_config* cfg);
// fix: 1 / 0 launches the iterations with / without the fix pass; -1 decides
// from the host mirror of the states.
int batch_run_iters(Ctx* c, Batch& b, int iters, bool use_graph, int fix = -1);
// Captures (or updates in place) the batch's graph of `iters` iterations
// without launching it.
int batch_graph_ready(Ctx* c, Batch& b, int iters, bool with_fix);
int batch_finalize(Ctx* c, Batch& b);
int read_states(Ctx* c, const Batch& b);
void fill_result(const Design& D, const DevState& s, ep_result* r);

}  // namespace ep
