// Host runtime of the MI355X placer: contexts, resident designs, batches,
// the device-driven iteration loop and the single-kernel evaluation calls.
//
// A resident design is uploaded once into one device arena (SoA, 256-byte
// aligned sub-buffers).  A batch is a list of resident designs plus the
// device descriptors, block prefixes and transform jobs its kernels need;
// one iteration of a batch is a fixed sequence of launches, captured once
// into a hipGraph and replayed, with the host only reading back the small
// FP64 state array between replays.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "ep_internal.h"
#include "ep_runtime.h"

namespace ep {

// ------------------------------------------------------------- errors ----

thread_local std::string g_tls_error;

int fail(Ctx* c, int code, const std::string& msg) {
    g_tls_error = msg;
    if (c) {
        std::lock_guard<std::mutex> g(c->err_mu);
        c->last_error = msg;
    }
    return code;
}

#define HIPCHK(ctx, expr)                                                             \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            return fail(ctx, EP_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

static size_t align_up(size_t v) { return (v + 255) & ~(size_t)255; }

// ------------------------------------------------------- transforms ------

// rustdct matrices (output index first): C2[k][j] = cos(pi k (2j+1) / 2n);
// C3[j][k] = k ? cos(pi k (2j+1) / 2n) : 1/2; S3[j][k] = sin(pi (k+1)(2j+1) / 2n)
// for k < n-1 and (-1)^j / 2 for k = n-1.  The IDXST mode replaces S3 by
// S3'[j][k] = -sin(pi k (2j+1) / 2n) (0 for k = 0).
static void build_matrix(int kind, int n, int mode, std::vector<double>& m) {
    m.assign((size_t)n * n, 0.0);
    const double pi = 3.14159265358979323846;
    for (int j = 0; j < n; j++)
        for (int k = 0; k < n; k++) {
            double v;
            if (kind == 0) v = std::cos(pi * j * (2.0 * k + 1.0) / (2.0 * n));  // C2[j][k]: j = freq
            else if (kind == 1) v = k == 0 ? 0.5 : std::cos(pi * k * (2.0 * j + 1.0) / (2.0 * n));
            else if (mode == kFieldReference)
                v = k < n - 1 ? std::sin(pi * (k + 1.0) * (2.0 * j + 1.0) / (2.0 * n))
                              : ((j & 1) ? -0.5 : 0.5);
            else
                v = k == 0 ? 0.0 : -std::sin(pi * k * (2.0 * j + 1.0) / (2.0 * n));
            m[(size_t)j * n + k] = v;
        }
}

static uint16_t bf16_bits(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    uint32_t r = u + 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(r >> 16);
}
static float bf16_float(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

int get_transform(Ctx* c, int n, int mode, Transform** out) {
    auto key = std::make_pair(n, mode);
    auto it = c->transforms.find(key);
    if (it != c->transforms.end()) {
        *out = it->second.get();
        return EP_OK;
    }
    std::unique_ptr<Transform> t(new Transform());
    t->n = n;
    t->mode = mode;
    const size_t nn = (size_t)n * n;
    const size_t bytes = 3 * align_up(nn * 8) + 3 * align_up(nn * 4) + 6 * align_up(nn * 2) +
                         3 * align_up((size_t)n * 8);
    HIPCHK(c, hipMalloc(&t->mem, bytes));
    char* p = (char*)t->mem;
    std::vector<double> m;
    std::vector<float> f(nn);
    std::vector<uint16_t> hi(nn), lo(nn);
    for (int kind = 0; kind < 3; kind++) {
        build_matrix(kind, n, mode, m);
        for (size_t i = 0; i < nn; i++) {
            f[i] = (float)m[i];
            hi[i] = bf16_bits(f[i]);
            lo[i] = bf16_bits(f[i] - bf16_float(hi[i]));
        }
        t->d[kind] = (double*)p; p += align_up(nn * 8);
        t->f[kind] = (float*)p; p += align_up(nn * 4);
        t->hi[kind] = p; p += align_up(nn * 2);
        t->lo[kind] = p; p += align_up(nn * 2);
        HIPCHK(c, hipMemcpy(t->d[kind], m.data(), nn * 8, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(t->f[kind], f.data(), nn * 4, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(t->hi[kind], hi.data(), nn * 2, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(t->lo[kind], lo.data(), nn * 2, hipMemcpyHostToDevice));
    }
    // FFT engine twiddles, computed in FP64 and rounded once
    {
        const double pi = 3.14159265358979323846;
        std::vector<float> tw(2 * (size_t)n), tws(2 * (size_t)n), dtw(2 * (size_t)n);
        fft_twiddle_table(n, false, tw.data());
        fft_twiddle_table(n, true, tws.data());
        for (int k = 0; k < n; k++) {
            dtw[2 * k] = (float)std::cos(pi * k / (2.0 * n));
            dtw[2 * k + 1] = (float)-std::sin(pi * k / (2.0 * n));
        }
        t->fft_tw = (float*)p; p += align_up((size_t)n * 8);
        t->fft_tw_solve = (float*)p; p += align_up((size_t)n * 8);
        t->dct_tw = (float*)p; p += align_up((size_t)n * 8);
        HIPCHK(c, hipMemcpy(t->fft_tw, tw.data(), (size_t)n * 8, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(t->fft_tw_solve, tws.data(), (size_t)n * 8, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(t->dct_tw, dtw.data(), (size_t)n * 8, hipMemcpyHostToDevice));
    }
    c->transform_bytes += bytes;
    *out = t.get();
    c->transforms[key] = std::move(t);
    return EP_OK;
}

// ------------------------------------------------------------ designs ----

// Net layout: nets of degree 2..64 go to bucket b (group size G = 2^(b+1),
// the smallest power of two >= degree); each net occupies G consecutive
// slots; every bucket segment is padded to whole waves.  Inside a bucket,
// nets are ordered by their smallest movable cell index, so nets of one
// cluster are neighbours in memory and so are the gradient slots a block of
// cells reads back.  Larger nets go to a CSR list handled one block per
// net; their pins take the slots after the wave slots.  Nets with fewer
// than two pins are skipped, as common.rs:242 does.  The cell -> slot CSR
// lists, for every movable cell, the slots of its pins in ascending order.
static void build_layout(const ep_design* d, NetLayout& L) {
    std::vector<std::vector<std::pair<uint32_t, int>>> nets(kMaxBuckets);
    L.hf_off.assign(1, 0);
    const uint32_t nm = (uint32_t)d->n_mov;
    std::vector<uint32_t> med;
    for (int e = 0; e < d->n_nets; e++) {
        uint32_t b0 = d->net_off[e], n = d->net_off[e + 1] - b0;
        if (n < 2) continue;
        if (n > 64) {
            for (uint32_t k = 0; k < n; k++) L.hf_pins.push_back(d->pin_cell[b0 + k]);
            L.hf_off.push_back((uint32_t)L.hf_pins.size());
            L.hf_net.push_back(e);
            continue;
        }
        int b = 0;
        while ((2u << b) < n) b++;
        // key: the median movable pin (a net's slots then sit beside the nets
        // of most of its cells, not beside one stray pin's)
        med.clear();
        for (uint32_t k = 0; k < n; k++)
            if (d->pin_cell[b0 + k] < nm) med.push_back(d->pin_cell[b0 + k]);
        uint32_t key = 0xFFFFFFFFu;
        if (!med.empty()) {
            std::nth_element(med.begin(), med.begin() + med.size() / 2, med.end());
            key = med[med.size() / 2];
        }
        nets[b].push_back({key, e});
    }
    int waves = 0;
    for (int b = 0; b < kMaxBuckets; b++) {
        std::stable_sort(nets[b].begin(), nets[b].end(),
                         [](const std::pair<uint32_t, int>& x, const std::pair<uint32_t, int>& y) {
                             return x.first < y.first;
                         });
        L.bucket_wave_start[b] = waves;
        L.bucket_group[b] = 2 << b;
        const uint32_t G = 2u << b;
        const size_t start = L.slots.size();
        for (auto& kv : nets[b]) {
            uint32_t b0 = d->net_off[kv.second], n = d->net_off[kv.second + 1] - b0;
            for (uint32_t k = 0; k < G; k++) {
                L.slots.push_back(k < n ? d->pin_cell[b0 + k] : kPad);
                L.slot_net.push_back(k < n ? kv.second : -1);
            }
        }
        while ((L.slots.size() - start) % kWave) {
            L.slots.push_back(kPad);
            L.slot_net.push_back(-1);
        }
        waves += (int)((L.slots.size() - start) / kWave);
    }
    L.bucket_wave_start[kMaxBuckets] = waves;
    // cell -> slot CSR over the slots of nets of <= 64 pins (counting sort;
    // slots visited in ascending order)
    const size_t nws = L.slots.size();
    L.cs_off.assign((size_t)nm + 1, 0);
    for (size_t s = 0; s < nws; s++) {
        uint32_t c = L.slots[s];
        if (c != kPad && c < nm) L.cs_off[c + 1]++;
    }
    for (size_t i = 0; i < nm; i++) L.cs_off[i + 1] += L.cs_off[i];
    L.cs_idx.resize(L.cs_off[nm]);
    std::vector<uint32_t> fillp(L.cs_off.begin(), L.cs_off.end() - 1);
    for (size_t s = 0; s < nws; s++) {
        uint32_t c = L.slots[s];
        if (c != kPad && c < nm) L.cs_idx[fillp[c]++] = (uint32_t)s;
    }
    // each cell's first kCellSlots slots at a fixed stride, the rest of the
    // row pointing at the zero slot past every real one (n_slots)
    const uint32_t zero_slot = (uint32_t)(L.slots.size() + L.hf_pins.size());
    L.cs8.assign((size_t)nm * kCellSlots, zero_slot);
    for (size_t i = 0; i < nm; i++)
        for (uint32_t q = 0; q < (uint32_t)kCellSlots && L.cs_off[i] + q < L.cs_off[i + 1]; q++)
            L.cs8[i * kCellSlots + q] = L.cs_idx[L.cs_off[i] + q];
    // cell -> high-fanout-net CSR (nets ascending, a net once per pin of the cell)
    L.chf_off.assign((size_t)nm + 1, 0);
    const size_t nhf = L.hf_off.size() - 1;
    for (size_t h = 0; h < nhf; h++)
        for (uint32_t p = L.hf_off[h]; p < L.hf_off[h + 1]; p++)
            if (L.hf_pins[p] < nm) L.chf_off[L.hf_pins[p] + 1]++;
    for (size_t i = 0; i < nm; i++) L.chf_off[i + 1] += L.chf_off[i];
    L.chf_net.resize(L.chf_off[nm]);
    std::vector<uint32_t> fillh(L.chf_off.begin(), L.chf_off.end() - 1);
    for (size_t h = 0; h < nhf; h++)
        for (uint32_t p = L.hf_off[h]; p < L.hf_off[h + 1]; p++)
            if (L.hf_pins[p] < nm) L.chf_net[fillh[L.hf_pins[p]]++] = (uint32_t)h;
}

static int validate_design(Ctx* c, const ep_design* d) {
    if (!d) return fail(c, EP_ERR_INVALID_ARG, "null design");
    if (d->n_mov < 0 || d->n_fixed < 0 || d->n_nets < 0 || d->n_pins < 0)
        return fail(c, EP_ERR_INVALID_ARG, "negative size");
    if (d->W < 1 || d->H < 1) return fail(c, EP_ERR_INVALID_ARG, "grid must be at least 1x1");
    int bx = d->Bx > 0 ? d->Bx : d->W, by = d->By > 0 ? d->By : d->H;
    if (bx < 1 || by < 1 || bx > 8192 || by > 8192)
        return fail(c, EP_ERR_INVALID_ARG, "bin grid out of range");
    if (d->n_mov > 0 && (!d->x0 || !d->y0)) return fail(c, EP_ERR_INVALID_ARG, "missing x0/y0");
    if (d->n_fixed > 0 && (!d->fixed_x || !d->fixed_y))
        return fail(c, EP_ERR_INVALID_ARG, "missing fixed coordinates");
    if (d->n_nets > 0 && !d->net_off) return fail(c, EP_ERR_INVALID_ARG, "missing net_off");
    if (d->n_nets > 0) {
        if (d->net_off[0] != 0 || (int64_t)d->net_off[d->n_nets] != d->n_pins)
            return fail(c, EP_ERR_INVALID_ARG, "net_off must start at 0 and end at n_pins");
        for (int e = 0; e < d->n_nets; e++)
            if (d->net_off[e + 1] < d->net_off[e])
                return fail(c, EP_ERR_INVALID_ARG, "net_off not monotone");
    }
    if (d->n_pins > 0 && !d->pin_cell) return fail(c, EP_ERR_INVALID_ARG, "missing pin_cell");
    const uint64_t ncell = (uint64_t)d->n_mov + (uint64_t)d->n_fixed;
    for (int64_t p = 0; p < d->n_pins; p++)
        if (d->pin_cell[p] >= ncell) return fail(c, EP_ERR_INVALID_ARG, "pin_cell out of range");
    for (int i = 0; i < d->n_mov; i++)
        if (!std::isfinite(d->x0[i]) || !std::isfinite(d->y0[i]))
            return fail(c, EP_ERR_NONFINITE, "non-finite initial position");
    if (d->area)
        for (int i = 0; i < d->n_mov; i++)
            if (!(d->area[i] > 0.0f) || !std::isfinite(d->area[i]))
                return fail(c, EP_ERR_INVALID_ARG, "area must be positive and finite");
    if (d->net_weight)
        for (int e = 0; e < d->n_nets; e++)
            if (!std::isfinite(d->net_weight[e]))
                return fail(c, EP_ERR_INVALID_ARG, "net weights must be finite");
    return EP_OK;
}

struct ArenaPlan {
    size_t off = 0;
    size_t take(size_t bytes) {
        size_t o = off;
        off += align_up(bytes > 0 ? bytes : 1);
        return o;
    }
};

// Computes the arena layout; with base == nullptr it only sizes it.
static size_t plan_arena(Design& D, char* base) {
    DesignDev& G = D.dev;
    const size_t T = D.prec == EP_PREC_FP64 ? 8 : 4;
    const size_t N = (size_t)G.n_mov, F = (size_t)G.n_fixed;
    const size_t B = (size_t)G.Bx * G.By;
    ArenaPlan a;
    auto P = [&](size_t bytes) -> void* { size_t o = a.take(bytes); return base ? base + o : nullptr; };
    for (int ax = 0; ax < 2; ax++) G.x0[ax] = P(N * T);
    for (int ax = 0; ax < 2; ax++) G.pos[ax] = P((N + F) * T);
    for (int ax = 0; ax < 2; ax++) G.dpos[ax] = P(N * T);
    G.look = P((N + F) * 2 * T);
    for (int p = 0; p < 2; p++)
        for (int ax = 0; ax < 2; ax++) G.g[p][ax] = P(N * T);
    G.pg = P(((size_t)G.n_slots + 1) * 2 * T);  // + the zero slot, never written
    G.cs8 = (uint32_t*)P(N * kCellSlots * 4);
    G.cs_off = (uint32_t*)P((N + 1) * 4);
    G.gwl = P(N * 2 * T);
    G.cs_idx = (uint32_t*)P(D.cs_count * 4);
    G.chf_off = (uint32_t*)P((N + 1) * 4);
    G.chf_net = (uint32_t*)P(D.chf_count * 4);
    G.hf_stat = P((size_t)G.n_hf * kHfStat * T);
    G.perm[0] = (uint32_t*)P(N * 4);
    G.perm[1] = (uint32_t*)P(N * 4);
    G.perm0 = (uint32_t*)P(N * 4);
    G.keys = (uint32_t*)P(N * 4);
    G.stage_xy = P(N * 2 * T);
    G.stage_hs = (float*)P(N * 4);
    G.sorted_xy = P(N * 2 * T);
    G.sorted_hs[0] = (float*)P(N * 4);
    G.sorted_hs[1] = (float*)P(N * 4);
    G.tcount = (uint32_t*)P((size_t)G.n_keys * 4);
    G.tstart = (uint32_t*)P(((size_t)G.n_keys + 1) * 4);
    G.tofs = (uint32_t*)P(N * 4);
    G.pinw = P(N * T);
    G.hs = (float*)P(N * 4);
    G.slot_w = D.weighted ? P((size_t)G.n_slots * T) : nullptr;
    G.slots = (uint32_t*)P((size_t)G.n_wa_waves * kWave * 4);
    G.hf_off = (uint32_t*)P((size_t)(G.n_hf + 1) * 4);
    G.hf_pins = (uint32_t*)P(D.hf_pin_count * 4);
    G.dacc = P(B * (D.prec == EP_PREC_FP64 ? 16 : 8));  // FP64: integer and fraction words
    G.dens = P(B * T);
    G.tT = P(B * T * G.ks_x);
    for (int ax = 0; ax < 2; ax++) G.ehat[ax] = P(B * T * G.ks_y);
    for (int ax = 0; ax < 2; ax++) G.uT[ax] = P(B * T * G.ks_x);
    G.field = P(B * 2 * T * G.ks_y);
    G.spec_x = P((size_t)2 * G.Bx * T);
    G.spec_y = P((size_t)2 * G.By * T);
    G.part_cell = (double*)P((size_t)G.n_cell_blocks * kCellParts * 8);
    G.part_bin = (double*)P((size_t)std::max(G.n_ov_parts, 1) * kBinParts * 8);
    G.part_net = (double*)P((size_t)(G.n_wa_waves + G.n_hf + 1) * 8);
    G.log = (ep_iter_rec*)P((size_t)D.log_cap * sizeof(ep_iter_rec));
    return a.off;
}

static void fill_desc_sizes(Design& D, const ep_design* d, const NetLayout& L) {
    DesignDev& G = D.dev;
    memset(&G, 0, sizeof(G));
    G.n_mov = d->n_mov;
    G.n_fixed = d->n_fixed;
    G.n_nets = d->n_nets;
    G.W = d->W;
    G.H = d->H;
    G.Bx = d->Bx > 0 ? d->Bx : d->W;
    G.By = d->By > 0 ? d->By : d->H;
    G.prec = D.prec;
    // K slices of the long transform products (FAST MFMA shapes, K >= 512):
    // slices of 128 keep each block's chain of K stages short
    const bool mfma_shape = D.prec == EP_PREC_FAST && G.Bx % 64 == 0 && G.By % 64 == 0;
    G.ks_x = mfma_shape && G.Bx >= kSplitMinK ? G.Bx / kSplitSlice : 1;
    G.ks_y = mfma_shape && G.By >= kSplitMinK ? G.By / kSplitSlice : 1;
    G.n_cell_blocks = (G.n_mov + kBlock - 1) / kBlock;
    G.n_scat_blocks = (G.n_mov + kScatCells - 1) / kScatCells;
    G.n_bin_blocks = (G.Bx * G.By + kBlock - 1) / kBlock;
    G.n_wa_waves = L.bucket_wave_start[kMaxBuckets];
    G.n_hf = (int)L.hf_off.size() - 1;
    G.n_buckets = kMaxBuckets;
    for (int b = 0; b <= kMaxBuckets; b++) G.bucket_wave_start[b] = L.bucket_wave_start[b];
    for (int b = 0; b < kMaxBuckets; b++) G.bucket_group[b] = L.bucket_group[b];
    G.rx = (double)G.Bx / (double)G.W;
    G.ry = (double)G.By / (double)G.H;
    G.max_x = (double)(G.W - 1);  // mod.rs:73-74
    G.max_y = (double)(G.H - 1);
    D.hf_pin_count = L.hf_pins.size();
    G.n_slots = G.n_wa_waves * kWave + (int)L.hf_pins.size();
    {
        G.tiles_x = (G.Bx + 7) / 8;
        G.tiles_y = (G.By + 7) / 8;
        const int tiles = std::max(G.tiles_x, G.tiles_y);
        int side = 1;
        while (side < tiles) side <<= 1;
        G.n_keys = side * side;
        // the pull scatter needs every footprint to span at most 9 bins per
        // axis (2 h r + 2 <= 9 with margin); wider cells use the atomic scatter
        double hs_max = 0.5;
        if (d->area)
            for (int i = 0; i < d->n_mov; i++)
                hs_max = std::max(hs_max, 0.5 * std::sqrt((double)d->area[i]));
        G.scat_pull = (2.0 * hs_max * std::max(G.rx, G.ry) <= 6.5) ? 1 : 0;
        if (G.scat_pull) {
            G.n_tile_blocks = ((G.tiles_x + 3) / 4) * ((G.tiles_y + 3) / 4);
            G.n_ov_parts = G.n_tile_blocks;
            G.n_scat_blocks = 0;
            G.n_bin_blocks = 0;
        } else {
            G.n_tile_blocks = 0;
            G.n_ov_parts = G.n_bin_blocks;
        }
    }
    D.cs_count = L.cs_off.empty() ? 0 : L.cs_off.back();
    D.chf_count = L.chf_off.empty() ? 0 : L.chf_off.back();
    D.weighted = d->net_weight != nullptr;
}

}  // namespace ep

using namespace ep;

extern "C" {

int ep_version(void) { return 1; }

void ep_config_default(ep_config* c) {
    if (!c) return;
    memset(c, 0, sizeof(*c));
    c->seed = 1;  // config.rs:27-38
    c->wl_coeff = 0.5;
    c->target_util = 0.7;
    c->timing_driven = 0;
    c->max_iters = 500;
    c->target_density = 1.0;
    c->timing_weight = 0.0;
    c->nesterov_step_size = 0.1;
    c->legalize_interval = 5;
    c->min_iters = 20;  // mod.rs:31-43
    c->overflow_target = 0.1;
    c->density_eta = 0.1;
    c->grow_mult = 1.025;
    c->grow_threshold = 50.0;
    c->grow_add = 1.0;
    c->stop_every_iter = 0;
    c->log_hpwl = 1;
    c->check_nonfinite = 1;
    c->field_mode = EP_FIELD_REFERENCE;
}

int ep_create(ep_ctx** out, int device, uint64_t budget) {
    if (!out) return fail(nullptr, EP_ERR_INVALID_ARG, "null out");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0)
        return fail(nullptr, EP_ERR_DEVICE, "no HIP device available (the placer has no CPU path)");
    if (device < 0 || device >= n) return fail(nullptr, EP_ERR_INVALID_ARG, "bad device index");
    std::unique_ptr<Ctx> c(new Ctx());
    c->device = device;
    HIPCHK(nullptr, hipSetDevice(device));
    size_t fr = 0, tot = 0;
    HIPCHK(nullptr, hipMemGetInfo(&fr, &tot));
    c->budget = budget ? budget : (uint64_t)(tot * 0.9);
    HIPCHK(nullptr, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIPCHK(nullptr, hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
    HIPCHK(nullptr, hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    HIPCHK(nullptr, hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
    HIPCHK(nullptr, hipEventCreateWithFlags(&c->ev_states, hipEventDisableTiming));
    {
        int khz = 0;
        HIPCHK(nullptr, hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device));
        c->wall_khz = khz > 0 ? (double)khz : 0.0;
    }
    HIPCHK(nullptr, hipMalloc(&c->state_pool, sizeof(DevState) * kStatePool));
    HIPCHK(nullptr, hipHostMalloc(&c->state_host, sizeof(DevState) * kStatePool, hipHostMallocDefault));
    for (int i = kStatePool - 1; i >= 0; i--) c->free_states.push_back(i);
    ep_config_default(&c->eval_cfg);
    c->t0 = now_ms();
    {
        int rc = noalloc_init(c.get());  // this device's ep_place_noalloc buffer
        if (rc) return rc;
    }
    *out = (ep_ctx*)c.release();
    return EP_OK;
}

void ep_destroy(ep_ctx* h) {
    Ctx* c = (Ctx*)h;
    if (!c) return;
    ep_serve_stop(h);
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    for (auto& kv : c->designs) {
        hipFree(kv.second->arena);
    }
    c->designs.clear();
    arena_pool_clear(c);
    for (auto& kv : c->transforms) hipFree(kv.second->mem);
    c->transforms.clear();
    release_batch_cache(c);
    hipFree(c->state_pool);
    hipHostFree(c->state_host);
    hipStreamSynchronize(c->side);
    hipEventDestroy(c->ev_fork);
    hipEventDestroy(c->ev_join);
    hipEventDestroy(c->ev_states);
    hipStreamDestroy(c->side);
    hipStreamDestroy(c->stream);
    delete c;
}

const char* ep_last_error(ep_ctx* h) {
    Ctx* c = (Ctx*)h;
    if (!c) return g_tls_error.c_str();
    std::lock_guard<std::mutex> g(c->err_mu);
    c->err_copy = c->last_error;
    return c->err_copy.c_str();
}

uint64_t ep_design_bytes(const ep_design* d, int32_t prec) {
    if (!d) return 0;
    NetLayout L;
    build_layout(d, L);
    Design D;
    D.prec = prec == EP_PREC_FP64 ? EP_PREC_FP64 : EP_PREC_FAST;
    D.log_cap = kDefaultLogCap;
    fill_desc_sizes(D, d, L);
    return (uint64_t)plan_arena(D, nullptr);
}

int ep_design_upload(ep_ctx* h, const ep_design* d, int32_t prec, uint32_t* id_out) {
    Ctx* c = (Ctx*)h;
    if (!c || !id_out) return fail(c, EP_ERR_INVALID_ARG, "null argument");
    if (prec != EP_PREC_FP64 && prec != EP_PREC_FAST)
        return fail(c, EP_ERR_INVALID_ARG, "unknown precision");
    Prepared P;
    int rc = prepare_design(c, d, prec, kDefaultLogCap, P);  // host work outside the lock
    if (rc) return rc;
    if (d->n_nets > 0) {
        P.D->net_off_h.assign(d->net_off, d->net_off + d->n_nets + 1);
        P.D->pin_cell_h.assign(d->pin_cell, d->pin_cell + d->n_pins);
    }
    std::lock_guard<std::recursive_mutex> g(c->mu);
    rc = commit_design(c, P, id_out);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));  // P's host vectors are read until here
    // the transform matrices of the design's bin shape, both field modes, so
    // a placement of it needs no allocation of its own (ep_place_noalloc)
    const DesignDev& G = c->designs[*id_out]->dev;
    for (int mode : {EP_FIELD_REFERENCE, EP_FIELD_IDXST}) {
        Transform* t = nullptr;
        rc = get_transform(c, G.Bx, mode, &t);
        if (!rc && G.By != G.Bx) rc = get_transform(c, G.By, mode, &t);
        if (rc) {
            free_design_locked(c, *id_out);
            return rc;
        }
    }
    return EP_OK;
}

}  // extern "C"

namespace ep {

template <typename T>
static void to_T(std::vector<char>& out, const double* p, size_t n) {
    out.resize(n * sizeof(T));
    T* o = (T*)out.data();
    for (size_t i = 0; i < n; i++) o[i] = (T)p[i];
}

static void to_prec(std::vector<char>& out, const double* p, size_t n, int prec) {
    if (prec == EP_PREC_FP64) to_T<double>(out, p, n);
    else to_T<float>(out, p, n);
}

int prepare_design(Ctx* c, const ep_design* d, int prec, int log_cap, Prepared& P) {
    int rc = validate_design(c, d);
    if (rc) return rc;
    build_layout(d, P.L);
    P.D.reset(new Design());
    Design& D = *P.D;
    D.prec = prec == EP_PREC_FP64 ? EP_PREC_FP64 : EP_PREC_FAST;
    D.log_cap = log_cap;
    fill_desc_sizes(D, d, P.L);
    P.bytes = plan_arena(D, nullptr);
    const size_t N = (size_t)d->n_mov, F = (size_t)d->n_fixed;
    to_prec(P.x0[0], d->x0, N, D.prec);
    to_prec(P.x0[1], d->y0, N, D.prec);
    D.fixed0[0].assign(d->fixed_x, d->fixed_x + F);
    D.fixed0[1].assign(d->fixed_y, d->fixed_y + F);
    to_prec(P.fixed[0], d->fixed_x, F, D.prec);
    to_prec(P.fixed[1], d->fixed_y, F, D.prec);
    {
        const size_t Tb = D.prec == EP_PREC_FP64 ? 8 : 4;
        P.fixed_xy.resize(2 * F * Tb);
        for (size_t k = 0; k < F; k++)
            for (int ax = 0; ax < 2; ax++) memcpy(&P.fixed_xy[(2 * k + ax) * Tb], &P.fixed[ax][k * Tb], Tb);
    }
    // pin weights, common.rs:381-407 (or the caller's)
    std::vector<double> w(N, 0.0);
    if (d->pinw) {
        w.assign(d->pinw, d->pinw + N);
    } else {
        for (int e = 0; e < d->n_nets; e++) {
            uint32_t b0 = d->net_off[e], n = d->net_off[e + 1] - b0;
            if (n < 2) continue;
            double wt = 1.0 / (double)n;
            for (uint32_t k = 0; k < n; k++) {
                uint32_t cc = d->pin_cell[b0 + k];
                if (cc < (uint32_t)d->n_mov) w[cc] += wt;
            }
        }
    }
    to_prec(P.pinw, w.data(), N, D.prec);
    P.hs.resize(N);
    for (size_t i = 0; i < N; i++)
        P.hs[i] = (float)(0.5 * std::sqrt((double)(d->area ? d->area[i] : 1.0f)));
    if (d->net_weight) {  // common.rs:246-249, per slot of the WA layout
        std::vector<double> sw((size_t)D.dev.n_slots, 0.0);
        for (size_t s = 0; s < P.L.slot_net.size(); s++)
            if (P.L.slot_net[s] >= 0) sw[s] = d->net_weight[P.L.slot_net[s]];
        const size_t base = P.L.slots.size();
        for (size_t h = 0; h + 1 < P.L.hf_off.size(); h++)
            for (uint32_t p = P.L.hf_off[h]; p < P.L.hf_off[h + 1]; p++)
                sw[base + p] = d->net_weight[P.L.hf_net[h]];
        to_prec(P.slot_w, sw.data(), sw.size(), D.prec);
    }
    // scatter order: any permutation of the cells is valid.  A placement starts
    // from the cells ordered by the (row-major) tile of their initial
    // position, so that the first sort, like every later one, walks cells
    // that mostly share a tile; in index order its waves saw dozens of keys
    P.ident.resize(N);
    for (size_t i = 0; i < N; i++) P.ident[i] = (uint32_t)i;
    {
        std::vector<uint64_t> key(N);
        for (size_t i = 0; i < N; i++) {
            const int tx = std::min(std::max((int)std::floor(d->x0[i] * D.dev.rx), 0), D.dev.Bx - 1) >> 3;
            const int ty = std::min(std::max((int)std::floor(d->y0[i] * D.dev.ry), 0), D.dev.By - 1) >> 3;
            key[i] = ((uint64_t)((uint32_t)ty * (uint32_t)D.dev.tiles_x + (uint32_t)tx) << 32) | i;
        }
        std::sort(key.begin(), key.end());
        P.order.resize(N);
        P.order_hs.resize(N);
        for (size_t j = 0; j < N; j++) {
            P.order[j] = (uint32_t)(key[j] & 0xffffffffu);
            P.order_hs[j] = P.hs[P.order[j]];
        }
    }
    const double pi = 3.14159265358979323846;
    for (int ax = 0; ax < 2; ax++) {  // density.rs:135-140
        const int B = ax == 0 ? D.dev.Bx : D.dev.By;
        std::vector<double> sp(2 * (size_t)B);
        for (int k = 0; k < B; k++) {
            double w0 = pi * k / (double)B;
            sp[2 * k] = w0;
            sp[2 * k + 1] = 2.0 * (1.0 - std::cos(w0));
        }
        to_prec(P.spec[ax], sp.data(), sp.size(), D.prec);
    }
    return EP_OK;
}

static int copy_in(Ctx* c, void* dst, const void* src, size_t n) {
    if (n) HIPCHK(c, hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, c->stream));
    return EP_OK;
}

int arena_alloc(Ctx* c, size_t bytes, void** out, size_t* got) {
    // a pooled arena of at least this size and at most a quarter (plus 1 MiB) larger
    auto it = c->arena_pool.lower_bound(bytes);
    if (it != c->arena_pool.end() && it->first <= bytes + bytes / 4 + (1u << 20)) {
        *out = it->second;
        *got = it->first;
        c->pooled -= it->first;
        c->arena_pool.erase(it);
        return EP_OK;
    }
    // make room: pooled arenas go back to the device, largest first
    while (!c->arena_pool.empty() && c->used + c->pooled + bytes > c->budget) {
        auto last = std::prev(c->arena_pool.end());
        hipFree(last->second);
        c->pooled -= last->first;
        c->arena_pool.erase(last);
    }
    hipError_t e = hipMalloc(out, bytes);
    if (e != hipSuccess) return fail(c, EP_ERR_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
    *got = bytes;
    return EP_OK;
}

void arena_release(Ctx* c, void* p, size_t bytes) {
    // keep up to a quarter of the budget (at most 16 GiB) for reuse
    const uint64_t cap = std::min<uint64_t>(c->budget / 4, 16ull << 30);
    if (c->pooled + bytes <= cap) {
        c->arena_pool.emplace(bytes, p);
        c->pooled += bytes;
    } else {
        hipFree(p);
    }
}

void arena_pool_clear(Ctx* c) {
    for (auto& kv : c->arena_pool) hipFree(kv.second);
    c->arena_pool.clear();
    c->pooled = 0;
}

int commit_design(Ctx* c, Prepared& P, uint32_t* id_out) {
    HIPCHK(c, hipSetDevice(c->device));
    Design& D = *P.D;
    if (c->used + P.bytes > c->budget) return fail(c, EP_ERR_OOM, "design does not fit in the HBM budget");
    if (c->free_states.empty()) return fail(c, EP_ERR_OOM, "state pool exhausted");
    size_t got = 0;
    int rc = arena_alloc(c, P.bytes, &D.arena, &got);
    if (rc) return rc;
    D.bytes = got;
    plan_arena(D, (char*)D.arena);
    const size_t N = (size_t)D.dev.n_mov;
    const NetLayout& L = P.L;
    hipError_t e = hipMemsetAsync(D.arena, 0, P.bytes, c->stream);
    if (e != hipSuccess) rc = fail(c, EP_ERR_DEVICE, std::string("hipMemsetAsync: ") + hipGetErrorString(e));
    // the fixed cells' locations follow the movable cells' entries of pos
    // and look (no kernel writes them), so a pin of any cell is one load
    const size_t Tb = D.prec == EP_PREC_FP64 ? 8 : 4;
    if (!rc) rc = copy_in(c, (char*)D.dev.look + 2 * N * Tb, P.fixed_xy.data(), P.fixed_xy.size());
    for (int ax = 0; ax < 2 && !rc; ax++) {
        rc = copy_in(c, D.dev.x0[ax], P.x0[ax].data(), P.x0[ax].size());
        if (!rc) rc = copy_in(c, (char*)D.dev.pos[ax] + N * Tb, P.fixed[ax].data(), P.fixed[ax].size());
        if (!rc) rc = copy_in(c, ax == 0 ? D.dev.spec_x : D.dev.spec_y, P.spec[ax].data(), P.spec[ax].size());
        if (!rc) rc = copy_in(c, D.dev.perm[ax], ax == 0 ? P.order.data() : P.ident.data(), N * 4);
        if (!rc) rc = copy_in(c, D.dev.sorted_hs[ax], ax == 0 ? P.order_hs.data() : P.hs.data(), N * 4);
    }
    if (!rc) rc = copy_in(c, D.dev.perm0, P.order.data(), N * 4);
    if (!rc) rc = copy_in(c, D.dev.pinw, P.pinw.data(), P.pinw.size());
    if (!rc) rc = copy_in(c, D.dev.hs, P.hs.data(), N * 4);
    if (!rc) rc = copy_in(c, D.dev.slots, L.slots.data(), L.slots.size() * 4);
    if (!rc) rc = copy_in(c, D.dev.hf_off, L.hf_off.data(), L.hf_off.size() * 4);
    if (!rc) rc = copy_in(c, D.dev.hf_pins, L.hf_pins.data(), L.hf_pins.size() * 4);
    if (!rc) rc = copy_in(c, D.dev.cs_off, L.cs_off.data(), L.cs_off.size() * 4);
    if (!rc) rc = copy_in(c, D.dev.cs_idx, L.cs_idx.data(), L.cs_idx.size() * 4);
    if (!rc) rc = copy_in(c, D.dev.cs8, L.cs8.data(), L.cs8.size() * 4);
    if (!rc) rc = copy_in(c, D.dev.chf_off, L.chf_off.data(), L.chf_off.size() * 4);
    if (!rc) rc = copy_in(c, D.dev.chf_net, L.chf_net.data(), L.chf_net.size() * 4);
    if (!rc && D.dev.slot_w) rc = copy_in(c, (void*)D.dev.slot_w, P.slot_w.data(), P.slot_w.size());
    if (rc) {
        hipStreamSynchronize(c->stream);
        arena_release(c, D.arena, D.bytes);
        D.arena = nullptr;
        return rc;
    }
    D.state_idx = c->free_states.back();
    c->free_states.pop_back();
    D.dev.st = c->state_pool + D.state_idx;
    D.id = c->next_id++;
    c->used += D.bytes;
    *id_out = D.id;
    c->designs[D.id] = std::move(P.D);
    return EP_OK;
}

int upload_locked(Ctx* c, const ep_design* d, int prec, int log_cap, uint32_t* id_out) {
    Prepared P;
    int rc = prepare_design(c, d, prec, log_cap, P);
    if (!rc) rc = commit_design(c, P, id_out);
    if (!rc) HIPCHK(c, hipStreamSynchronize(c->stream));
    return rc;
}

}  // namespace ep
