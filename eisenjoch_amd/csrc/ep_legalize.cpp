// Greedy legalization with the reference's exact selection rule
// (crates/nextpnr/src/placer/electro_place/mod.rs:206-256): cells in solver
// index order each take the free BEL of their bucket with the smallest
// (x - bx)^2 + (y - by)^2, ties going to the BEL that comes first in bucket
// order (BEL index order, crates/nextpnr/src/context/buckets.rs:37-43); each
// root's cluster children are placed right after it (place_cluster_children,
// crates/nextpnr/src/placer/common.rs:319-374).
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

#include "ep_legal.h"
#include "eplace.h"

namespace ep {

Legalizer::Legalizer(int32_t n_bels, const int32_t* bel_x, const int32_t* bel_y, const int32_t* bel_bucket)
    : n_bels_(n_bels), bx_(bel_x), by_(bel_y), bb_(bel_bucket) {
    std::map<int32_t, std::vector<int32_t>> members;
    for (int32_t b = 0; b < n_bels; b++) members[bel_bucket[b]].push_back(b);
    for (auto& kv : members) {
        BucketGrid& G = grids_[kv.first];
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
        for (int32_t b : kv.second) {  // ascending b keeps bucket order inside a location
            size_t t = (size_t)(bel_y[b] - ymn) * G.w + (bel_x[b] - xmn);
            G.bels[fillp[t]++] = b;
        }
        G.order = kv.second;
    }
}

// The free BEL of G nearest to (cx, cy), first minimum in bucket order; -1 if none.
int Legalizer::nearest(const BucketGrid& G, double cx, double cy, const uint8_t* taken, double* cost) const {
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
            if (taken[b]) continue;
            const double dx = (double)bx_[b] - cx;
            const double dy = (double)by_[b] - cy;
            const double c = dx * dx + dy * dy;
            if (c < best || (c == best && b < bb)) {
                best = c;
                bb = b;
            }
        }
    };
    for (int r = 0; r <= rmax; r++) {
        // every BEL of ring r lies at least r - e from the cell
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
    *cost = best;
    return bb;
}

int Legalizer::run(int32_t n_cells, const double* x, const double* y, const int32_t* cell_bucket,
                   const ep_legal_problem* lp, uint8_t* taken, int32_t* cell_bel, int32_t* child_bel,
                   double* displacement) const {
    const int32_t nch = lp ? lp->n_children : 0;
    // first-free cursor per bucket for the children's fallback: BELs are only
    // ever taken during one run, so a cursor that skips taken BELs never
    // passes a free one
    std::map<int32_t, size_t> cursor;
    int32_t ch = 0;
    double disp = 0.0;
    for (int32_t i = 0; i < n_cells; i++) {
        auto it = grids_.find(cell_bucket[i]);
        if (it == grids_.end()) return EP_ERR_NO_BELS;
        const double cx = x[i], cy = y[i];
        if (!std::isfinite(cx) || !std::isfinite(cy)) return EP_ERR_NONFINITE;
        double cost = 0.0;
        const int32_t b = nearest(it->second, cx, cy, taken, &cost);
        if (b < 0) return EP_ERR_NO_BELS;
        taken[b] = 1;
        cell_bel[i] = b;
        disp += cost;
        // place_cluster_children: exact location first, then the first free BEL
        for (; ch < nch && lp->child_root[ch] == i; ch++) {
            auto gt = grids_.find(lp->child_bucket[ch]);
            int32_t cb = -1;
            if (gt != grids_.end()) {
                const BucketGrid& G = gt->second;
                const int tx = bx_[b] + lp->child_dx[ch], ty = by_[b] + lp->child_dy[ch];
                if (tx >= G.x0 && ty >= G.y0 && tx < G.x0 + G.w && ty < G.y0 + G.h) {
                    const size_t t = (size_t)(ty - G.y0) * G.w + (tx - G.x0);
                    for (int32_t k = G.start[t]; k < G.start[t + 1] && cb < 0; k++)
                        if (!taken[G.bels[k]]) cb = G.bels[k];
                }
                if (cb < 0) {
                    size_t& cur = cursor[lp->child_bucket[ch]];
                    while (cur < G.order.size() && taken[G.order[cur]]) cur++;
                    if (cur < G.order.size()) cb = G.order[cur];
                }
            }
            if (cb < 0) return EP_ERR_PLACEMENT_FAILED;
            taken[cb] = 1;
            child_bel[ch] = cb;
        }
    }
    if (displacement) *displacement = disp;
    return EP_OK;
}

int check_legal_problem(const ep_legal_problem* lp, int32_t n_mov, int32_t n_fixed, int32_t n_nets,
                        const char** why) {
    auto bad = [&](const char* m) { *why = m; return EP_ERR_INVALID_ARG; };
    if (!lp) return bad("null legalization problem");
    if (lp->n_bels < 0 || lp->n_children < 0) return bad("negative size");
    if (lp->n_bels > 0 && (!lp->bel_x || !lp->bel_y || !lp->bel_bucket)) return bad("missing BEL arrays");
    if (n_mov > 0 && !lp->cell_bucket) return bad("missing cell_bucket");
    if (lp->n_children > 0 && (!lp->child_root || !lp->child_dx || !lp->child_dy || !lp->child_bucket ||
                               !lp->child_fixed))
        return bad("missing cluster child arrays");
    for (int32_t k = 0; k < lp->n_children; k++) {
        if (lp->child_root[k] < 0 || lp->child_root[k] >= n_mov) return bad("child_root out of range");
        if (k > 0 && lp->child_root[k] < lp->child_root[k - 1]) return bad("cluster children not grouped by ascending root");
        if (lp->child_fixed[k] < -1 || lp->child_fixed[k] >= n_fixed) return bad("child_fixed out of range");
    }
    (void)n_nets;
    return EP_OK;
}

double legal_hpwl(const ep_design* d, const ep_legal_problem* lp, const int32_t* cell_bel,
                  const int32_t* child_bel) {
    // the location of every fixed cell: its own, or its cluster child's BEL
    std::vector<int32_t> fx((size_t)d->n_fixed), fy((size_t)d->n_fixed);
    for (int32_t j = 0; j < d->n_fixed; j++) {
        fx[j] = (int32_t)d->fixed_x[j];
        fy[j] = (int32_t)d->fixed_y[j];
    }
    for (int32_t k = 0; k < lp->n_children; k++)
        if (lp->child_fixed[k] >= 0) {
            fx[lp->child_fixed[k]] = lp->bel_x[child_bel[k]];
            fy[lp->child_fixed[k]] = lp->bel_y[child_bel[k]];
        }
    double total = 0.0;
    for (int32_t e = 0; e < d->n_nets; e++) {
        const uint32_t b0 = d->net_off[e], n = d->net_off[e + 1] - b0;
        if (n < 2) continue;  // no users, or nothing placed but one pin
        if (lp->net_has_driver && !lp->net_has_driver[e]) continue;  // wirelength.rs:16-18
        int32_t xmn = INT32_MAX, xmx = INT32_MIN, ymn = INT32_MAX, ymx = INT32_MIN;
        for (uint32_t k = 0; k < n; k++) {
            const uint32_t c = d->pin_cell[b0 + k];
            int32_t px, py;
            if (c < (uint32_t)d->n_mov) {
                px = lp->bel_x[cell_bel[c]];
                py = lp->bel_y[cell_bel[c]];
            } else {
                px = fx[c - d->n_mov];
                py = fy[c - d->n_mov];
            }
            xmn = std::min(xmn, px); xmx = std::max(xmx, px);
            ymn = std::min(ymn, py); ymx = std::max(ymx, py);
        }
        total += (double)((xmx - xmn) + (ymx - ymn));
    }
    return total;
}

}  // namespace ep

using namespace ep;

extern "C" int ep_legalize(int32_t n_cells, const double* x, const double* y, const int32_t* cell_bucket,
                           int32_t n_bels, const int32_t* bel_x, const int32_t* bel_y,
                           const int32_t* bel_bucket, uint8_t* bel_taken, int32_t* cell_bel,
                           double* displacement) {
    if (n_cells < 0 || n_bels < 0) return EP_ERR_INVALID_ARG;
    if (n_cells > 0 && (!x || !y || !cell_bucket || !cell_bel)) return EP_ERR_INVALID_ARG;
    if (n_bels > 0 && (!bel_x || !bel_y || !bel_bucket || !bel_taken)) return EP_ERR_INVALID_ARG;
    Legalizer L(n_bels, bel_x, bel_y, bel_bucket);
    return L.run(n_cells, x, y, cell_bucket, nullptr, bel_taken, cell_bel, nullptr, displacement);
}

extern "C" int ep_legalize_problem(const ep_legal_problem* lp, int32_t n_cells, const double* x,
                                   const double* y, int32_t* cell_bel, int32_t* child_bel,
                                   double* displacement) {
    const char* why = nullptr;
    if (check_legal_problem(lp, n_cells, INT32_MAX, 0, &why)) return EP_ERR_INVALID_ARG;
    if (n_cells > 0 && (!x || !y || !cell_bel)) return EP_ERR_INVALID_ARG;
    if (lp->n_children > 0 && !child_bel) return EP_ERR_INVALID_ARG;
    std::vector<uint8_t> taken((size_t)lp->n_bels, 0);
    if (lp->bel_taken) taken.assign(lp->bel_taken, lp->bel_taken + lp->n_bels);
    Legalizer L(lp->n_bels, lp->bel_x, lp->bel_y, lp->bel_bucket);
    return L.run(n_cells, x, y, lp->cell_bucket, lp, taken.data(), cell_bel, child_bel, displacement);
}

extern "C" int ep_legal_hpwl(const ep_design* d, const ep_legal_problem* lp, const int32_t* cell_bel,
                             const int32_t* child_bel, double* hpwl) {
    if (!d || !lp || !hpwl || (d->n_mov > 0 && !cell_bel)) return EP_ERR_INVALID_ARG;
    if (d->n_nets > 0 && (!d->net_off || !d->pin_cell)) return EP_ERR_INVALID_ARG;
    if (d->n_fixed > 0 && (!d->fixed_x || !d->fixed_y)) return EP_ERR_INVALID_ARG;
    const char* why = nullptr;
    if (check_legal_problem(lp, d->n_mov, d->n_fixed, d->n_nets, &why)) return EP_ERR_INVALID_ARG;
    if (lp->n_children > 0 && !child_bel) return EP_ERR_INVALID_ARG;
    for (int32_t i = 0; i < d->n_mov; i++)
        if (cell_bel[i] < 0 || cell_bel[i] >= lp->n_bels) return EP_ERR_INVALID_ARG;
    for (int32_t k = 0; k < lp->n_children; k++)
        if (child_bel[k] < 0 || child_bel[k] >= lp->n_bels) return EP_ERR_INVALID_ARG;
    *hpwl = legal_hpwl(d, lp, cell_bel, child_bel);
    return EP_OK;
}
