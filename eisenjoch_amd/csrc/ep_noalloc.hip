// ep_place_noalloc: the placement of one resident design without any device
// allocation, for callers that cannot allocate device memory (serving loops
// that must not hit the allocator, harnesses that forbid it).
//
// ep_place builds a batch image per new design list (per-design descriptors,
// block prefixes and tables, transform jobs) in a fresh device allocation and
// instantiates a hipGraph for its chunks of iterations.  This variant writes
// the same image into `scratch`, a file-scope device array of kScratchBytes
// that every context on a device shares, and launches the iterations eagerly
// (the same kernels in the same order, so the results are the ep_place ones
// bit for bit).  Everything else the loop touches was allocated before: the
// design arena and its transform matrices at upload, the state pool, streams
// and events at ep_create.
//
// The array's host name is only a stand-in: host code uses the device address
// that hipGetSymbolAddress returns, looked up once per device with that
// device current (ep_create calls noalloc_init) and kept per device.  Each
// call rewrites the whole image before its first launch, so nothing of an
// earlier call's contents is read.  Calls that overlap on one device (two
// contexts, two threads) race for the array; calls through one context are
// serialised by its lock.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "ep_internal.h"
#include "ep_runtime.h"

namespace ep {

// 1 MiB holds the image of a batch of one design up to tens of millions of
// pins (about 4 bytes per 256 cells, per 768 net slots and per bin tile).
constexpr size_t kScratchBytes = 1u << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

// Looks up this device's copy of `scratch` (device must be current).
int noalloc_init(Ctx* c) {
    if (c->device < 0 || c->device >= kMaxDevices)
        return fail(c, EP_ERR_INVALID_ARG, "device index beyond the allocation-free table");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (g_scratch_ptr[c->device]) return EP_OK;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(c, EP_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[c->device] = p;
    return EP_OK;
}

static void* scratch_ptr(int device) {
    std::lock_guard<std::mutex> g(g_scratch_mu);
    return (device >= 0 && device < kMaxDevices) ? g_scratch_ptr[device] : nullptr;
}

}  // namespace ep

using namespace ep;

extern "C" int ep_place_noalloc(ep_ctx* h, uint32_t id, const ep_config* cfg, ep_result* res) {
    Ctx* c = (Ctx*)h;
    if (!c || !cfg || !res) return fail(c, EP_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::recursive_mutex> g(c->mu);
    if (c->server_running) return fail(c, EP_ERR_INVALID_ARG, "context is serving; stop it first");
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return fail(c, EP_ERR_DEVICE, hipGetErrorString(e));
    void* mem = scratch_ptr(c->device);
    if (!mem) return fail(c, EP_ERR_DEVICE, "allocation-free buffer not set up on this device");
    Design* D = find_design(c, id);
    if (!D) return fail(c, EP_ERR_NOT_FOUND, "unknown design id");
    int rc = validate_config(c, cfg, D->log_cap);
    if (rc) return rc;
    // the transform matrices of the design's shape were built at upload; a
    // missing one would need an allocation
    if (!c->transforms.count({D->dev.Bx, cfg->field_mode}) || !c->transforms.count({D->dev.By, cfg->field_mode}))
        return fail(c, EP_ERR_INVALID_ARG, "transform matrices of this bin shape are not prepared");
    Batch b;
    std::vector<Design*> ds{D};
    std::vector<const ep_config*> cf{cfg};
    rc = batch_build_cfgs(c, ds, cf, b, mem, kScratchBytes);
    if (rc) return rc;
    const int sync_every = cfg->legalize_interval > 0 ? cfg->legalize_interval : 5;
    rc = batch_place(c, b, sync_every, false);
    if (!rc) {
        fill_result(*D, c->state_host[D->state_idx], res);
        res->t_submit_ms = 0.0;
        res->t_admit_ms = 0.0;
        res->t_done_ms = b.done_ms.empty() ? 0.0 : b.done_ms[0];
        if (res->status != EP_OK) rc = fail(c, res->status, "non-finite value in the placement state");
    }
    batch_release(b);  // nothing to free: the image lives in the shared array
    return rc;
}
