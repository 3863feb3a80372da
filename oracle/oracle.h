/*
 * ElectroPlace oracle: a plain-C FP64 restatement of the reference's
 * electrostatic placer (crates/nextpnr/src/placer/electro_place/ and
 * crates/nextpnr/src/placer/solver/{wa,nesterov}.rs).
 *
 * TEST INFRASTRUCTURE ONLY.  Only the test-suite, the smoke check and the
 * CPU-baseline leg of the benchmark load this library, and only as the
 * checker.  The product library never links or calls it.
 *
 * Every function cites the reference file:line it restates.  The one
 * extension (cell area and bins != tiles) reduces exactly to the reference
 * when every area is 1 and Bx == W, By == H.
 */
#ifndef EP_ORACLE_H
#define EP_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Flattened design.  Cells 0..n_mov-1 are movable (solver index order),
 * pin ids >= n_mov index fixed cells (id - n_mov).  Each net's pins are the
 * driver first and then the users, as common.rs:228-240 gathers them. */
typedef struct orc_design {
    int32_t n_mov, n_fixed, n_nets;
    int32_t W, H, Bx, By;
    int64_t n_pins;
    const double* x0;
    const double* y0;
    const float* area;       /* NULL => unit area (the reference) */
    const double* fixed_x;
    const double* fixed_y;
    const uint32_t* net_off; /* n_nets + 1 */
    const uint32_t* pin_cell;/* n_pins */
    const double* pinw;      /* NULL => derived from the CSR degrees */
    const double* net_weight;/* NULL => every net weighs 1 (ElectroPlace passes None, mod.rs:114) */
} orc_design;

typedef struct orc_cfg {
    uint64_t seed;
    double wl_coeff;
    double target_density;
    double nesterov_step_size;
    int32_t max_iters;
    int32_t legalize_interval;
    int32_t stop_every_iter;   /* 0: reference rule (check on legalize iters) */
    int32_t min_iters;         /* reference: 20 */
    double overflow_target;    /* reference: 0.1 */
    int32_t field_mode;        /* 0: reference field; 1: index-shifted (IDXST) field */
} orc_cfg;

typedef struct orc_iter_rec {
    int32_t iter;
    int32_t checked;           /* 1 on a convergence-check iteration */
    double overflow;
    double lambda;
    double step_size_x, step_size_y;
    double step_norm_x, step_norm_y;
    double hpwl;               /* continuous HPWL of x after the step (check iters) */
    double wl_norm, den_norm;
} orc_iter_rec;

typedef struct orc_result {
    int32_t iters;             /* iterations executed */
    int32_t converged;
    double overflow;           /* overflow of the last executed iteration */
    double hpwl;               /* continuous HPWL of the final positions */
    double lambda;
} orc_result;

/* ---- weighted-average wirelength, wa.rs ---- */
double orc_wa_axis_value(const double* c, int n, double a);
void orc_wa_axis_grad(const double* c, int n, double a, double* g);
double orc_wa_wirelength(const double* px, const double* py, int n, double a);
void orc_wa_gradient(const orc_design* d, const double* x, const double* y,
                     double a, double* gx, double* gy);
void orc_pin_weights(const orc_design* d, double* pinw);

/* ---- density, density.rs ---- */
void orc_bin_range(const orc_design* d, double cx, double cy, double area,
                   int32_t out[4], double box[4]);
void orc_density(const orc_design* d, const double* x, const double* y, double* dens);
double orc_overlap(const double* dens, int64_t n);
void orc_dct2(double* v, int n);
void orc_dct3(double* v, int n);
void orc_dst3(double* v, int n);
void orc_dct2_naive(const double* in, double* out, int n);
void orc_dct3_naive(const double* in, double* out, int n);
void orc_dst3_naive(const double* in, double* out, int n);
void orc_field(const double* dens, int Bx, int By, double target_density,
               double* ex, double* ey);
void orc_field_mode(const double* dens, int Bx, int By, double target_density, int mode,
                    double* ex, double* ey);
void orc_density_gradient(const orc_design* d, const double* x, const double* y,
                          const double* ex, const double* ey, double* gx, double* gy);

/* ---- Nesterov solver, nesterov.rs ---- */
typedef struct orc_nesterov {
    int n;
    double* x;
    double* x_prev;
    double step_size;
    int iter;
    double a;
} orc_nesterov;
void orc_nesterov_init(orc_nesterov* s, int n, double step, double* xbuf, double* xpbuf);
void orc_nesterov_set_positions(orc_nesterov* s, const double* x);
double orc_nesterov_momentum(const orc_nesterov* s);
void orc_nesterov_look_ahead(const orc_nesterov* s, double* y);
double orc_nesterov_step(orc_nesterov* s, const double* g);
void orc_nesterov_clamp(orc_nesterov* s, double lo, double hi);
int orc_nesterov_bb(const orc_nesterov* s, const double* gp, const double* g, double* out);
double orc_nesterov_lipschitz(const orc_nesterov* s, const double* gp, const double* g);
void orc_nesterov_adaptive_restart(orc_nesterov* s, const double* g);

/* ---- metrics ---- */
double orc_hpwl(const orc_design* d, const double* x, const double* y);

/* ---- the loop, mod.rs:65-204 (continuous part) ---- */
int orc_place(const orc_design* d, const orc_cfg* cfg, double* x_out, double* y_out,
              orc_iter_rec* log, int log_cap, int* n_log, orc_result* res);

/* ---- RNG, rng.rs ---- */
uint64_t orc_rng_next(uint64_t* state);
void orc_rng_shuffle_u32(uint64_t* state, uint32_t* v, int n);

/* ---- greedy legalization (mod.rs:206-256) and legalized HPWL (wirelength.rs:10-51) ---- */
int orc_legalize(int n_cells, const double* x, const double* y, const int32_t* cell_bucket,
                 int n_bels, const int32_t* bel_x, const int32_t* bel_y,
                 const int32_t* bel_bucket, uint8_t* bel_taken, int32_t* cell_bel,
                 double* displacement);
double orc_total_hpwl_int(const orc_design* d, const int32_t* cell_ix, const int32_t* cell_iy);

/* ---- place_electro end to end: legalization with cluster children, the
 *      legalized HPWL and the best snapshot (mod.rs:176-199,
 *      common.rs:319-374, 485-495, wirelength.rs:10-51) ---- */
typedef struct orc_legal {
    int32_t n_bels, n_children;
    const int32_t* bel_x;
    const int32_t* bel_y;
    const int32_t* bel_bucket;
    const uint8_t* bel_taken;     /* NULL: none bound on entry */
    const int32_t* cell_bucket;
    const int32_t* child_root;    /* ascending; a root's children in constraint order */
    const int32_t* child_dx;
    const int32_t* child_dy;
    const int32_t* child_bucket;
    const int32_t* child_fixed;   /* fixed-cell index of a child with pins, or -1 */
    const uint8_t* net_has_driver;/* NULL: every net has one */
} orc_legal;
int orc_legalize_full(const orc_design* d, const orc_legal* lp, const double* x, const double* y,
                      uint8_t* taken, int32_t* cell_bel, int32_t* child_bel, double* displacement);
double orc_legal_hpwl(const orc_design* d, const orc_legal* lp, const int32_t* cell_bel,
                      const int32_t* child_bel);
int orc_place_legalized(const orc_design* d, const orc_cfg* cfg, const orc_legal* lp,
                        int32_t* cell_bel, int32_t* child_bel, double* best_hpwl, int* best_iter,
                        int* legalizations, orc_result* res);

#ifdef __cplusplus
}
#endif
#endif
