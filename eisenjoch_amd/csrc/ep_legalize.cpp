// Greedy legalization with the reference's exact selection rule
// (crates/nextpnr/src/placer/electro_place/mod.rs:206-256): cells in solver
// index order each take the free BEL of their bucket with the smallest
// (x - bx)^2 + (y - by)^2, ties going to the BEL that comes first in bucket
// order (BEL index order, crates/nextpnr/src/context/buckets.rs:37-43).
//
// The reference scans the whole bucket per cell, O(N * |bucket|).  Here each
// bucket's BELs are binned by integer location and searched in square rings
// around the cell; a ring is only opened while its distance lower bound can
// still beat or tie the best cost found, so the result is identical to the
// full scan.
#include <algorithm>
#include <cmath>
#include <limits>
#include <map>
#include <vector>

#include "eplace.h"

namespace {

struct BucketGrid {
    int x0 = 0, y0 = 0, w = 0, h = 0;
    std::vector<int32_t> start;  // tile -> first index into bels (CSR)
    std::vector<int32_t> bels;   // BEL indices, ascending within a tile
};

}  // namespace

extern "C" int ep_legalize(int32_t n_cells, const double* x, const double* y, const int32_t* cell_bucket,
                           int32_t n_bels, const int32_t* bel_x, const int32_t* bel_y,
                           const int32_t* bel_bucket, uint8_t* bel_taken, int32_t* cell_bel,
                           double* displacement) {
    if (n_cells < 0 || n_bels < 0) return EP_ERR_INVALID_ARG;
    if (n_cells > 0 && (!x || !y || !cell_bucket || !cell_bel)) return EP_ERR_INVALID_ARG;
    if (n_bels > 0 && (!bel_x || !bel_y || !bel_bucket || !bel_taken)) return EP_ERR_INVALID_ARG;
    std::map<int32_t, BucketGrid> grids;
    {
        std::map<int32_t, std::vector<int32_t>> members;
        for (int32_t b = 0; b < n_bels; b++) members[bel_bucket[b]].push_back(b);
        for (auto& kv : members) {
            BucketGrid& G = grids[kv.first];
            int xmn = INT32_MAX, xmx = INT32_MIN, ymn = INT32_MAX, ymx = INT32_MIN;
            for (int32_t b : kv.second) {
                xmn = std::min(xmn, bel_x[b]); xmx = std::max(xmx, bel_x[b]);
                ymn = std::min(ymn, bel_y[b]); ymx = std::max(ymx, bel_y[b]);
            }
            G.x0 = xmn; G.y0 = ymn; G.w = xmx - xmn + 1; G.h = ymx - ymn + 1;
            std::vector<int32_t> cnt((size_t)G.w * G.h + 1, 0);
            for (int32_t b : kv.second) cnt[(size_t)(bel_y[b] - ymn) * G.w + (bel_x[b] - xmn) + 1]++;
            for (size_t t = 1; t < cnt.size(); t++) cnt[t] += cnt[t - 1];
            G.start = cnt;
            G.bels.resize(kv.second.size());
            std::vector<int32_t> fillp(cnt.begin(), cnt.end() - 1);
            for (int32_t b : kv.second) {  // ascending b keeps bucket order inside a tile
                size_t t = (size_t)(bel_y[b] - ymn) * G.w + (bel_x[b] - xmn);
                G.bels[fillp[t]++] = b;
            }
        }
    }
    double disp = 0.0;
    for (int32_t i = 0; i < n_cells; i++) {
        auto it = grids.find(cell_bucket[i]);
        if (it == grids.end()) return EP_ERR_NO_BELS;
        const BucketGrid& G = it->second;
        const double cx = x[i], cy = y[i];
        if (!std::isfinite(cx) || !std::isfinite(cy)) return EP_ERR_NONFINITE;
        const int tx0 = (int)std::min<double>(std::max<double>(std::floor(cx + 0.5), G.x0), G.x0 + G.w - 1);
        const int ty0 = (int)std::min<double>(std::max<double>(std::floor(cy + 0.5), G.y0), G.y0 + G.h - 1);
        const double e = std::max(std::fabs(tx0 - cx), std::fabs(ty0 - cy));
        const int rmax = std::max(std::max(tx0 - G.x0, G.x0 + G.w - 1 - tx0),
                                  std::max(ty0 - G.y0, G.y0 + G.h - 1 - ty0));
        double best = std::numeric_limits<double>::infinity();
        int32_t bb = -1;
        auto visit = [&](int tx, int ty) {
            if (tx < G.x0 || ty < G.y0 || tx >= G.x0 + G.w || ty >= G.y0 + G.h) return;
            const size_t t = (size_t)(ty - G.y0) * G.w + (tx - G.x0);
            for (int32_t k = G.start[t]; k < G.start[t + 1]; k++) {
                const int32_t b = G.bels[k];
                if (bel_taken[b]) continue;
                const double dx = (double)bel_x[b] - cx;
                const double dy = (double)bel_y[b] - cy;
                const double cost = dx * dx + dy * dy;
                if (cost < best || (cost == best && b < bb)) {
                    best = cost;
                    bb = b;
                }
            }
        };
        for (int r = 0; r <= rmax; r++) {
            const double lb = r - e - 1e-9;
            if (bb >= 0 && lb > 0.0 && lb * lb > best) break;
            if (r == 0) {
                visit(tx0, ty0);
                continue;
            }
            for (int tx = tx0 - r; tx <= tx0 + r; tx++) {
                visit(tx, ty0 - r);
                visit(tx, ty0 + r);
            }
            for (int ty = ty0 - r + 1; ty <= ty0 + r - 1; ty++) {
                visit(tx0 - r, ty);
                visit(tx0 + r, ty);
            }
        }
        if (bb < 0) return EP_ERR_NO_BELS;
        bel_taken[bb] = 1;
        cell_bel[i] = bb;
        disp += best;
    }
    if (displacement) *displacement = disp;
    return EP_OK;
}
