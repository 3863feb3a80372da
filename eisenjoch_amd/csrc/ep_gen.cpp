// Synthetic generic-FPGA design generator.
//
// Builds clustered-locality netlists on a W x H tile grid: movable cells are
// split into contiguous clusters, each net draws its degree from the
// 60/25/12/3 % mix (2, 3-4, 5-16, 17-64 pins, uniform inside a bucket) and
// takes `locality` of its pins from one cluster.  Fixed IO cells sit evenly
// spaced on the perimeter and drive evenly spaced nets.  Initial positions
// follow the reference's initial_placement (crates/nextpnr/src/placer/
// common.rs:27-117): per cell type, the BELs of its bucket are shuffled with
// the reference xorshift64 (crates/nextpnr/src/context/rng.rs:11-61, seed 0
// taken as 1) and the cells bind to them in index order.  The netlist itself
// uses a separate splitmix64 stream so the two never interact.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "eplace.h"

namespace {

struct SplitMix {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    // uniform integer in [0, n)
    uint32_t below(uint32_t n) { return (uint32_t)(((next() >> 32) * (uint64_t)n) >> 32); }
    double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

// rng.rs:26-31 and rng.rs:52-61
struct XorShift {
    uint64_t s;
    uint64_t next() {
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        return s;
    }
    void shuffle(std::vector<int32_t>& v) {
        size_t n = v.size();
        if (n <= 1) return;
        for (size_t i = n - 1; i >= 1; i--) {
            uint32_t j = ((uint32_t)next()) % (uint32_t)(i + 1);
            std::swap(v[i], v[j]);
        }
    }
};

int draw_degree(SplitMix& r) {
    double u = r.unit();
    if (u < 0.60) return 2;
    if (u < 0.85) return 3 + (int)r.below(2);
    if (u < 0.97) return 5 + (int)r.below(12);
    return 17 + (int)r.below(48);
}

}  // namespace

extern "C" int ep_generate(const ep_gen_params* p, ep_gen_design** out) {
    if (!p || !out) return EP_ERR_INVALID_ARG;
    *out = nullptr;
    const int N = p->n_cells, F = p->n_fixed, W = p->W, H = p->H;
    const int T = p->n_types < 1 ? 1 : p->n_types;
    if (N < 0 || F < 0 || p->n_nets < 0 || W < 1 || H < 1 || T > 2) return EP_ERR_INVALID_ARG;
    if ((int64_t)N > (int64_t)W * H) return EP_ERR_NO_BELS;
    const int C = p->n_clusters < 1 ? 1 : p->n_clusters;
    const int E = p->n_nets + (p->n_high_fanout > 0 ? p->n_high_fanout : 0);

    ep_gen_design* g = (ep_gen_design*)calloc(1, sizeof(ep_gen_design));
    if (!g) return EP_ERR_OOM;
    SplitMix rng{p->seed * 0x2545F4914F6CDD1Dull + 0x1234567ull};

    // cell types: type 1 spread evenly with fraction type1_frac
    int32_t* ctype = (int32_t*)malloc(sizeof(int32_t) * (size_t)(N > 0 ? N : 1));
    for (int i = 0; i < N; i++) {
        int t = 0;
        if (T == 2) {
            double f = p->type1_frac;
            t = (std::floor((i + 1) * f) > std::floor(i * f)) ? 1 : 0;
        }
        ctype[i] = t;
    }
    float* area = nullptr;
    if (p->area_mix) {
        area = (float*)malloc(sizeof(float) * (size_t)(N > 0 ? N : 1));
        for (int i = 0; i < N; i++) {
            double u = rng.unit();
            area[i] = u < 0.85 ? 1.0f : (u < 0.97 ? 2.0f : 4.0f);
        }
    }

    // IO on the perimeter, evenly spaced along the boundary walk
    double* fx = (double*)malloc(sizeof(double) * (size_t)(F > 0 ? F : 1));
    double* fy = (double*)malloc(sizeof(double) * (size_t)(F > 0 ? F : 1));
    {
        std::vector<std::pair<int, int>> ring;
        for (int x = 0; x < W; x++) ring.push_back({x, 0});
        for (int y = 1; y < H; y++) ring.push_back({W - 1, y});
        if (H > 1)
            for (int x = W - 2; x >= 0; x--) ring.push_back({x, H - 1});
        if (W > 1)
            for (int y = H - 2; y >= 1; y--) ring.push_back({0, y});
        for (int j = 0; j < F; j++) {
            size_t k = (size_t)((double)j * ring.size() / (double)F);
            fx[j] = ring[k].first;
            fy[j] = ring[k].second;
        }
    }

    // nets
    std::vector<uint32_t> off;
    std::vector<uint32_t> pins;
    off.reserve((size_t)E + 1);
    pins.reserve((size_t)E * 5);
    off.push_back(0);
    const int csz = N > 0 ? (N + C - 1) / C : 1;
    std::vector<uint32_t> tmp;
    for (int e = 0; e < E; e++) {
        int deg;
        if (e < p->n_nets) deg = draw_degree(rng);
        else deg = p->hf_min + (int)rng.below((uint32_t)(p->hf_max - p->hf_min + 1));
        if (N == 0) deg = 0;
        int cl = (int)rng.below((uint32_t)C);
        int c0 = cl * csz, c1 = c0 + csz < N ? c0 + csz : N;
        if (c1 <= c0) { c0 = 0; c1 = N; }
        tmp.clear();
        for (int k = 0; k < deg; k++) {
            uint32_t c = 0;
            for (int tries = 0; tries < 16; tries++) {
                if (rng.unit() < p->locality) c = (uint32_t)(c0 + (int)rng.below((uint32_t)(c1 - c0)));
                else c = rng.below((uint32_t)N);
                bool dup = false;
                if (deg <= 64)
                    for (uint32_t q : tmp) dup |= (q == c);
                if (!dup) break;
            }
            tmp.push_back(c);
        }
        pins.insert(pins.end(), tmp.begin(), tmp.end());
        off.push_back((uint32_t)pins.size());
    }
    // IO j drives net floor(j * n_nets / F): its pin replaces the driver
    for (int j = 0; j < F && p->n_nets > 0; j++) {
        int e = (int)((int64_t)j * p->n_nets / F);
        if (off[e + 1] > off[e]) pins[off[e]] = (uint32_t)(N + j);
    }

    // logic BELs: one per type per tile, tile-major (y outer, x inner)
    const int nb = W * H * T;
    int32_t* bx = (int32_t*)malloc(sizeof(int32_t) * (size_t)nb);
    int32_t* by = (int32_t*)malloc(sizeof(int32_t) * (size_t)nb);
    int32_t* bt = (int32_t*)malloc(sizeof(int32_t) * (size_t)nb);
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++)
            for (int t = 0; t < T; t++) {
                int b = (y * W + x) * T + t;
                bx[b] = x; by[b] = y; bt[b] = t;
            }
    // initial_placement (common.rs:79-113): per type, shuffle the bucket and bind in order
    int32_t* ibel = (int32_t*)malloc(sizeof(int32_t) * (size_t)(N > 0 ? N : 1));
    double* x0 = (double*)malloc(sizeof(double) * (size_t)(N > 0 ? N : 1));
    double* y0 = (double*)malloc(sizeof(double) * (size_t)(N > 0 ? N : 1));
    XorShift xs{p->seed == 0 ? 1ull : p->seed};
    for (int t = 0; t < T; t++) {
        std::vector<int32_t> avail;
        avail.reserve((size_t)W * H);
        for (int b = 0; b < nb; b++)
            if (bt[b] == t) avail.push_back(b);
        std::vector<int32_t> cells;
        for (int i = 0; i < N; i++)
            if (ctype[i] == t) cells.push_back(i);
        if (cells.empty()) continue;
        xs.shuffle(avail);
        if (cells.size() > avail.size()) {
            free(ctype); free(area); free(fx); free(fy); free(bx); free(by); free(bt);
            free(ibel); free(x0); free(y0); free(g);
            return EP_ERR_NO_BELS;
        }
        for (size_t k = 0; k < cells.size(); k++) {
            int b = avail[k];
            ibel[cells[k]] = b;
            x0[cells[k]] = bx[b];
            y0[cells[k]] = by[b];
        }
    }

    uint32_t* noff = (uint32_t*)malloc(sizeof(uint32_t) * off.size());
    memcpy(noff, off.data(), sizeof(uint32_t) * off.size());
    uint32_t* npin = (uint32_t*)malloc(sizeof(uint32_t) * (pins.size() > 0 ? pins.size() : 1));
    if (!pins.empty()) memcpy(npin, pins.data(), sizeof(uint32_t) * pins.size());

    g->d.n_mov = N;
    g->d.n_fixed = F;
    g->d.n_nets = E;
    g->d.W = W;
    g->d.H = H;
    g->d.Bx = p->Bx > 0 ? p->Bx : W;
    g->d.By = p->By > 0 ? p->By : H;
    g->d.n_pins = (int64_t)pins.size();
    g->d.x0 = x0;
    g->d.y0 = y0;
    g->d.area = area;
    g->d.fixed_x = fx;
    g->d.fixed_y = fy;
    g->d.net_off = noff;
    g->d.pin_cell = npin;
    g->d.pinw = nullptr;
    g->n_types = T;
    g->n_bels = nb;
    g->cell_type = ctype;
    g->bel_x = bx;
    g->bel_y = by;
    g->bel_type = bt;
    g->init_bel = ibel;
    *out = g;
    return EP_OK;
}

extern "C" void ep_gen_free(ep_gen_design* g) {
    if (!g) return;
    free((void*)g->d.x0);
    free((void*)g->d.y0);
    free((void*)g->d.area);
    free((void*)g->d.fixed_x);
    free((void*)g->d.fixed_y);
    free((void*)g->d.net_off);
    free((void*)g->d.pin_cell);
    free(g->cell_type);
    free(g->bel_x);
    free(g->bel_y);
    free(g->bel_type);
    free(g->init_bel);
    free(g);
}
