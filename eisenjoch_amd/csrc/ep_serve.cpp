// Serving: continuous batching of placement requests on one GPU.
//
// Requests are queued by ep_serve_submit (the design is copied), admitted in
// arrival order while a slot is free and the design's arena fits in the HBM
// budget, and then advance inside the running batch: every fused iteration
// launch covers all admitted designs, each at its own iteration count.  A
// design leaves the batch when its own state says it is done; its slot and
// memory are reused by the next admission.  Because the FAST path's sums are
// fixed-point and all reductions are per design in a fixed order, a design's
// trajectory does not depend on which other designs share its launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "ep_internal.h"
#include "ep_runtime.h"

namespace ep {


enum ReqState { kQueued = 0, kRunning = 1, kDone = 2 };

struct ServeReq {
    uint64_t ticket = 0;
    std::unique_ptr<Prepared> prep;  // built on the submitting thread; kept until done
    int n_mov = 0;
    ep_config cfg{};
    int state = kQueued;
    int status = EP_PENDING;
    bool cancel = false;
    ep_result res{};
    std::vector<double> xo, yo;
    Design* D = nullptr;
    uint64_t bytes = 0;
};

struct Server {
    std::thread th;
    std::mutex m;
    std::condition_variable cv;
    bool stop = false;
    ep_serve_config scfg{};
    std::deque<ServeReq*> queue;
    std::vector<ServeReq*> active;
    std::map<uint64_t, std::unique_ptr<ServeReq>> all;
    uint64_t next_ticket = 1;
    Batch batch;
    bool dirty = true;
    int failed = EP_OK;        // status of the failure that stopped the serving thread
    std::string failed_msg;
    int64_t fail_after = -1;   // EP_SERVE_FAIL_AFTER: chunks before a forced failure (tests)
    ep_serve_stats stats{};
};

void ServerDeleter::operator()(Server* s) const { delete s; }

static BatchDev single_view(const Batch& b, int i) {
    BatchDev one = b.bd;
    one.d = b.d_dev + i;
    one.nd = 1;
    one.pfx_cell = one.pfx_bin = one.pfx_wa = one.pfx_hf = one.pfx_scat = one.pfx_tile = b.dzero;
    one.blk_cell = one.blk_bin = one.blk_wa = one.blk_hf = one.blk_scat = one.blk_tile = nullptr;
    return one;
}

// Reduce pending log partials, compute the final HPWL and read the state of
// design i of the batch, leaving the others untouched.
static int finalize_one(Ctx* c, Batch& b, int i) {
    const Design& D = *b.ds[i];
    BatchDev one = single_view(b, i);
    launch_ctrl(one, 2, c->stream);
    launch_hpwl(one, b.prec, wa_blocks_of(D.dev.n_wa_waves), D.dev.n_hf, 1, c->stream);
    launch_ctrl(one, 3, c->stream);
    hipError_t e = hipMemcpyAsync(&c->state_host[D.state_idx], D.dev.st, sizeof(DevState),
                                  hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(c, EP_ERR_DEVICE, hipGetErrorString(e));
    return EP_OK;
}

static int copy_positions(Ctx* c, const Design& D, std::vector<double>& x, std::vector<double>& y) {
    const size_t N = (size_t)D.dev.n_mov;
    x.resize(N);
    y.resize(N);
    if (!N) return EP_OK;
    for (int ax = 0; ax < 2; ax++) {
        double* out = ax ? y.data() : x.data();
        hipError_t e;
        if (D.prec == EP_PREC_FP64) {
            e = hipMemcpy(out, D.dev.pos[ax], N * 8, hipMemcpyDeviceToHost);
        } else {
            std::vector<float> f(N);
            e = hipMemcpy(f.data(), D.dev.pos[ax], N * 4, hipMemcpyDeviceToHost);
            for (size_t k = 0; k < N; k++) out[k] = f[k];
        }
        if (e != hipSuccess) return fail(c, EP_ERR_DEVICE, hipGetErrorString(e));
    }
    return EP_OK;
}

// Ends every queued and running request with status rc (after a device
// failure the serving thread stops; the designs' memory is released).
static void fail_all(Ctx* c, Server& S, int rc) {
    std::string msg;
    {
        std::lock_guard<std::mutex> eg(c->err_mu);
        msg = c->last_error;
    }
    std::vector<ServeReq*> live;
    {
        std::lock_guard<std::mutex> lk(S.m);
        S.failed = rc;
        S.failed_msg = msg;
        for (ServeReq* r : S.active)
            if (r) live.push_back(r);
        for (ServeReq* r : S.queue) live.push_back(r);
        S.active.clear();
        S.queue.clear();
    }
    for (ServeReq* r : live) {
        if (r->D) free_design_locked(c, r->D->id);
        std::lock_guard<std::mutex> lk(S.m);
        r->D = nullptr;
        r->prep.reset();
        r->res.status = rc;
        r->res.t_done_ms = now_ms() - c->t0;
        r->status = rc;
        r->state = kDone;
    }
    batch_release(S.batch);
}

static void serve_loop(Ctx* c) {
    Server& S = *c->server;
    hipSetDevice(c->device);
    std::vector<int> fresh;  // batch indices admitted since the last rebuild
    std::vector<ServeReq*> fresh_reqs;
    int rc = EP_OK;
    for (;;) {
        std::vector<ServeReq*> admit;
        {
            std::unique_lock<std::mutex> lk(S.m);
            if (S.stop) break;
            // drop cancelled requests still in the queue
            for (auto it = S.queue.begin(); it != S.queue.end();) {
                if ((*it)->cancel) {
                    (*it)->state = kDone;
                    (*it)->status = EP_ERR_CANCELLED;
                    (*it)->res.status = EP_ERR_CANCELLED;
                    (*it)->res.t_done_ms = now_ms() - c->t0;
                    it = S.queue.erase(it);
                } else {
                    ++it;
                }
            }
            uint64_t free_b = c->budget > c->used ? c->budget - c->used : 0;
            uint64_t pending = 0;
            while (!S.queue.empty() && S.active.size() + admit.size() < S.scfg.max_slots) {
                ServeReq* r = S.queue.front();
                if (pending + r->bytes > free_b) break;  // arrival order: no overtaking
                pending += r->bytes;
                admit.push_back(r);
                S.queue.pop_front();
            }
            if (admit.empty() && S.active.empty()) {
                S.cv.wait_for(lk, std::chrono::milliseconds(2));
                continue;
            }
        }
        std::lock_guard<std::recursive_mutex> g(c->mu);
        for (ServeReq* r : admit) {
            uint32_t id = 0;
            int arc = commit_design(c, *r->prep, &id);
            std::lock_guard<std::mutex> lk(S.m);
            r->res.t_admit_ms = now_ms() - c->t0;
            if (arc) {  // this request alone (budget, state pool)
                r->state = kDone;
                r->status = arc;
                r->res.status = arc;
                r->res.t_done_ms = r->res.t_admit_ms;
                continue;
            }
            r->D = c->designs[id].get();
            r->state = kRunning;
            S.active.push_back(r);
            fresh_reqs.push_back(r);
            S.dirty = true;
            S.stats.admitted++;
        }
        if (S.active.empty()) continue;
        if (S.dirty) {
            S.batch.keep_graphs = getenv("EP_SERVE_NO_GRAPH_UPDATE") == nullptr;
            std::vector<Design*> ds;
            std::vector<const ep_config*> cf;
            for (ServeReq* r : S.active) {
                ds.push_back(r->D);
                cf.push_back(&r->cfg);
            }
            if ((rc = batch_build_cfgs(c, ds, cf, S.batch, nullptr, 0))) break;
            S.stats.batch_builds++;
            fresh.clear();
            for (size_t i = 0; i < S.active.size(); i++)
                if (std::find(fresh_reqs.begin(), fresh_reqs.end(), S.active[i]) != fresh_reqs.end())
                    fresh.push_back((int)i);
            if (!fresh.empty() && (rc = batch_reset_some(c, S.batch, fresh))) break;
            fresh_reqs.clear();
            S.dirty = false;
        }
        if (S.fail_after >= 0 && (int64_t)S.stats.chunks >= S.fail_after) {
            // forced from the host before any launch of this chunk (tests of the failure path)
            rc = fail(c, EP_ERR_DEVICE, "forced serving failure (EP_SERVE_FAIL_AFTER)");
            break;
        }
        const int caps = c->graph_captures, ups = c->graph_updates;
        const bool graphs = getenv("EP_NO_GRAPH") == nullptr;
        if (graphs && S.batch.graph_stale[0] && S.batch.keep_graphs) {
            // the first chunk of a new membership runs eagerly, and the graph of
            // the next chunks (new designs have their penalty by then) is
            // recorded and updated while the GPU runs it
            if ((rc = batch_run_iters(c, S.batch, S.scfg.iters_per_sync, false))) break;
            if ((rc = batch_graph_ready(c, S.batch, S.scfg.iters_per_sync, false))) break;
        } else if ((rc = batch_run_iters(c, S.batch, S.scfg.iters_per_sync, graphs))) {
            break;
        }
        S.stats.graph_captures += (uint64_t)(c->graph_captures - caps);
        S.stats.graph_updates += (uint64_t)(c->graph_updates - ups);
        S.stats.chunks++;
        if ((rc = read_states(c, S.batch))) break;
        bool changed = false;
        for (int i = 0; i < (int)S.active.size() && !rc; i++) {
            ServeReq* r = S.active[i];
            const DevState& st = c->state_host[r->D->state_idx];
            bool cancel;
            {
                std::lock_guard<std::mutex> lk(S.m);
                cancel = r->cancel;
            }
            if (!st.done && !cancel) continue;
            std::vector<double> xo, yo;
            if ((rc = finalize_one(c, S.batch, i))) break;
            if ((rc = copy_positions(c, *r->D, xo, yo))) break;
            ep_result res = r->res;
            fill_result(*r->D, c->state_host[r->D->state_idx], &res);
            if (cancel) res.status = EP_ERR_CANCELLED;
            res.t_done_ms = now_ms() - c->t0;
            free_design_locked(c, r->D->id);
            r->prep.reset();  // its copies completed at the last synchronisation
            std::lock_guard<std::mutex> lk(S.m);
            r->D = nullptr;
            r->xo.swap(xo);
            r->yo.swap(yo);
            r->res = res;
            r->status = res.status;
            r->state = kDone;
            S.active[i] = nullptr;
            S.stats.finished++;
            changed = true;
        }
        if (rc) break;
        if (changed) {
            std::lock_guard<std::mutex> lk(S.m);
            S.active.erase(std::remove(S.active.begin(), S.active.end(), nullptr), S.active.end());
            S.dirty = true;
            if (S.active.empty()) {
                // nothing left to run: free the launch descriptors; the graphs
                // stay (stale) for the next admission to update in place
                if (S.batch.keep_graphs) {
                    std::vector<Design*> none;
                    std::vector<const ep_config*> nocfg;
                    batch_build_cfgs(c, none, nocfg, S.batch, nullptr, 0);
                } else {
                    batch_release(S.batch);
                }
            }
        }
    }
    std::lock_guard<std::recursive_mutex> g(c->mu);
    if (rc) fail_all(c, S, rc);
    batch_release(S.batch);
}

}  // namespace ep

using namespace ep;

extern "C" {

int ep_serve_start(ep_ctx* h, const ep_serve_config* scfg) {
    Ctx* c = (Ctx*)h;
    if (!c) return fail(nullptr, EP_ERR_INVALID_ARG, "null context");
    std::lock_guard<std::recursive_mutex> g(c->mu);
    if (c->server_running) return fail(c, EP_ERR_INVALID_ARG, "already serving");
    release_batch_cache(c);
    c->server.reset(new Server());
    ep_serve_config s{128, 5, 0};
    if (scfg) s = *scfg;
    if (s.max_slots == 0 || s.max_slots > 128) return fail(c, EP_ERR_INVALID_ARG, "max_slots must be 1..128");
    if (s.iters_per_sync == 0) s.iters_per_sync = 5;
    if (s.hbm_budget_bytes) c->budget = std::min<uint64_t>(c->budget, s.hbm_budget_bytes);
    c->server->scfg = s;
    if (const char* fa = getenv("EP_SERVE_FAIL_AFTER")) c->server->fail_after = strtoll(fa, nullptr, 10);
    c->server_running = true;
    c->server->th = std::thread(serve_loop, c);
    return EP_OK;
}

int ep_serve_stop(ep_ctx* h) {
    Ctx* c = (Ctx*)h;
    if (!c || !c->server) return EP_OK;
    {
        std::lock_guard<std::mutex> lk(c->server->m);
        c->server->stop = true;
    }
    c->server->cv.notify_all();
    if (c->server->th.joinable()) c->server->th.join();
    std::lock_guard<std::recursive_mutex> g(c->mu);
    for (ServeReq* r : c->server->active)
        if (r && r->D) free_design_locked(c, r->D->id);
    c->server.reset();
    c->server_running = false;
    return EP_OK;
}

int ep_serve_submit(ep_ctx* h, const ep_design* d, const ep_config* cfg, uint64_t* ticket) {
    Ctx* c = (Ctx*)h;
    if (!c || !d || !cfg || !ticket) return fail(c, EP_ERR_INVALID_ARG, "null argument");
    if (!c->server_running) return fail(c, EP_ERR_INVALID_ARG, "serving not started");
    int rc = validate_config(c, cfg, 1 << 20);
    if (rc) return rc;
    std::unique_ptr<ServeReq> r(new ServeReq());
    // layout, sizes and initial contents on the caller's thread, so the
    // serving thread only allocates and queues copies
    r->prep.reset(new Prepared());
    rc = prepare_design(c, d, EP_PREC_FAST, std::max(1, cfg->max_iters), *r->prep);
    if (rc) return rc;
    r->n_mov = d->n_mov;
    r->cfg = *cfg;
    r->bytes = r->prep->bytes;
    if (r->bytes > c->budget) return fail(c, EP_ERR_ADMISSION_REFUSED, "design exceeds the HBM budget");
    Server& S = *c->server;
    std::lock_guard<std::mutex> lk(S.m);
    if (S.failed) return fail(c, S.failed, "serving stopped after a failure: " + S.failed_msg);
    r->ticket = S.next_ticket++;
    r->res.status = EP_PENDING;
    r->res.t_submit_ms = now_ms() - c->t0;
    *ticket = r->ticket;
    S.queue.push_back(r.get());
    S.all[r->ticket] = std::move(r);
    S.cv.notify_all();
    return EP_OK;
}

int ep_serve_poll(ep_ctx* h, uint64_t ticket, ep_result* res) {
    Ctx* c = (Ctx*)h;
    if (!c || !c->server) return fail(c, EP_ERR_INVALID_ARG, "serving not started");
    Server& S = *c->server;
    std::lock_guard<std::mutex> lk(S.m);
    auto it = S.all.find(ticket);
    if (it == S.all.end()) return fail(c, EP_ERR_NOT_FOUND, "unknown ticket");
    if (res) *res = it->second->res;
    if (it->second->state != kDone) return EP_PENDING;
    return it->second->status;
}

int ep_serve_positions(ep_ctx* h, uint64_t ticket, double* x, double* y) {
    Ctx* c = (Ctx*)h;
    if (!c || !c->server || !x || !y) return fail(c, EP_ERR_INVALID_ARG, "null argument");
    Server& S = *c->server;
    std::lock_guard<std::mutex> lk(S.m);
    auto it = S.all.find(ticket);
    if (it == S.all.end()) return fail(c, EP_ERR_NOT_FOUND, "unknown ticket");
    if (it->second->state != kDone) return EP_PENDING;
    if (it->second->xo.empty() && it->second->n_mov > 0)
        return fail(c, it->second->status ? it->second->status : EP_ERR_NOT_FOUND, "no positions");
    std::copy(it->second->xo.begin(), it->second->xo.end(), x);
    std::copy(it->second->yo.begin(), it->second->yo.end(), y);
    return EP_OK;
}

int ep_serve_release(ep_ctx* h, uint64_t ticket) {
    Ctx* c = (Ctx*)h;
    if (!c || !c->server) return fail(c, EP_ERR_INVALID_ARG, "serving not started");
    Server& S = *c->server;
    std::lock_guard<std::mutex> lk(S.m);
    auto it = S.all.find(ticket);
    if (it == S.all.end()) return fail(c, EP_ERR_NOT_FOUND, "unknown ticket");
    if (it->second->state != kDone) return fail(c, EP_ERR_INVALID_ARG, "request still pending");
    S.all.erase(it);
    return EP_OK;
}

int ep_serve_stats_get(ep_ctx* h, ep_serve_stats* out) {
    Ctx* c = (Ctx*)h;
    if (!c || !out) return fail(c, EP_ERR_INVALID_ARG, "null argument");
    if (!c->server) return fail(c, EP_ERR_INVALID_ARG, "serving not started");
    Server& S = *c->server;
    std::lock_guard<std::mutex> lk(S.m);
    *out = S.stats;
    out->failed = S.failed;
    return EP_OK;
}

int ep_serve_cancel(ep_ctx* h, uint64_t ticket) {
    Ctx* c = (Ctx*)h;
    if (!c || !c->server) return fail(c, EP_ERR_INVALID_ARG, "serving not started");
    Server& S = *c->server;
    std::lock_guard<std::mutex> lk(S.m);
    auto it = S.all.find(ticket);
    if (it == S.all.end()) return fail(c, EP_ERR_NOT_FOUND, "unknown ticket");
    it->second->cancel = true;
    S.cv.notify_all();
    return EP_OK;
}

}  // extern "C"
