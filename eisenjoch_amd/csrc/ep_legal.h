// Host legalization shared by ep_legalize, ep_legalize_problem and the
// best-snapshot placement (ep_place_legalized).
#pragma once
#include <stdint.h>

#include <map>
#include <vector>

#include "eplace.h"

namespace ep {

// One bucket's BELs binned by integer location (CSR over the bucket's
// bounding box, BEL indices ascending inside a location) plus the bucket's
// BELs in bucket order (BEL index order).
struct BucketGrid {
    int x0 = 0, y0 = 0, w = 0, h = 0;
    std::vector<int32_t> start;
    std::vector<int32_t> bels;
    std::vector<int32_t> order;  // every BEL of the bucket, ascending
};

class Legalizer {
  public:
    // Builds the per-bucket grids (BEL arrays must outlive the object).
    Legalizer(int32_t n_bels, const int32_t* bel_x, const int32_t* bel_y, const int32_t* bel_bucket);
    // mod.rs:206-256 for the cells in solver order, with each root's cluster
    // children placed right after it (common.rs:319-374).  `taken` holds the
    // BELs bound on entry and is updated.  Returns EP_OK, EP_ERR_NO_BELS,
    // EP_ERR_PLACEMENT_FAILED or EP_ERR_NONFINITE.
    int run(int32_t n_cells, const double* x, const double* y, const int32_t* cell_bucket,
            const ep_legal_problem* lp, uint8_t* taken, int32_t* cell_bel, int32_t* child_bel,
            double* displacement) const;

  private:
    int nearest(const BucketGrid& G, double cx, double cy, const uint8_t* taken, double* cost) const;
    int32_t n_bels_;
    const int32_t *bx_, *by_, *bb_;
    std::map<int32_t, BucketGrid> grids_;
};

// Checks a legalization view against a design of n_mov movable and n_fixed
// fixed cells and n_nets nets (message in *why).
int check_legal_problem(const ep_legal_problem* lp, int32_t n_mov, int32_t n_fixed, int32_t n_nets,
                        const char** why);

// wirelength.rs:10-51 on a legalized binding: movable cells at their BELs,
// cluster children with pins at theirs, other fixed cells at their locations.
double legal_hpwl(const ep_design* d, const ep_legal_problem* lp, const int32_t* cell_bel,
                  const int32_t* child_bel);

}  // namespace ep
