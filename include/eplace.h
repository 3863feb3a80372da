/*
 * eplace.h -- C ABI of the MI355X electrostatic global placer.
 *
 * This library is the device half of the reference's ElectroPlace placer
 * (crates/nextpnr/src/placer/electro_place/mod.rs:65-204, `place_electro`).
 * The reference selects it by the string "electro" in the PyO3 `place()`
 * match (crates/npnr-python/src/lib.rs:293-302) and through the `Placer`
 * trait (crates/nextpnr/src/placer/mod.rs:31-52).  A host that owns the
 * netlist flattens it into an `ep_design` (solver-index order, nets in slot
 * order, driver pin first) and calls `ep_place`; legalization and BEL
 * binding stay on the host (`ep_legalize` offers the reference's greedy
 * rule).  Plain pointers and sizes only; every buffer is caller-owned and
 * copied.  All entry points are thread-safe per context.
 */
#ifndef EPLACE_H
#define EPLACE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  -2 and -3 mirror PlacerError::NoBelsAvailable and
 * PlacerError::PlacementFailed (crates/nextpnr/src/placer/mod.rs:20-28). */
#define EP_OK 0
#define EP_PENDING 1
#define EP_ERR_INVALID_ARG (-1)
#define EP_ERR_NO_BELS (-2)
#define EP_ERR_PLACEMENT_FAILED (-3)
#define EP_ERR_OOM (-4)
#define EP_ERR_ADMISSION_REFUSED (-5)
#define EP_ERR_DEVICE (-6)
#define EP_ERR_NONFINITE (-7)
#define EP_ERR_NOT_FOUND (-8)
#define EP_ERR_CANCELLED (-9)

/* Arithmetic of a resident design. */
#define EP_PREC_FP64 0 /* plain placer: FP64 everywhere, FP64 dense transforms */
#define EP_PREC_FAST 1 /* FP32 bulk, BF16 hi+lo MFMA transforms, FP64 reductions */

/* Field of the spectral solve.  REFERENCE feeds the DCT-indexed spectrum
 * straight into DST-III exactly as density.rs:155-163 does (the sine of
 * frequency k+1 carries coefficient k).  IDXST shifts it by one and negates
 * it, the textbook field -d(phi)/dx whose force spreads cells; it is the only
 * departure from the reference and exists because the reference field does
 * not reach overflow <= 0.10 on the synthetic designs. */
#define EP_FIELD_REFERENCE 0
#define EP_FIELD_IDXST 1

typedef struct ep_ctx ep_ctx;

/* A flattened placement problem.  Movable cells are 0..n_mov-1 in solver
 * index order (common.rs:155-174); pin ids >= n_mov name fixed cells
 * (id - n_mov).  Each net's pins are its collected pins, driver first
 * (common.rs:228-240).  Bx/By = 0 means bins = tiles (the reference,
 * mod.rs:75-76).  area = NULL means unit cells (the reference); pinw =
 * NULL derives common.rs:381-407 weights from the CSR degrees. */
typedef struct ep_design {
    int32_t n_mov;
    int32_t n_fixed;
    int32_t n_nets;
    int32_t W, H;
    int32_t Bx, By;
    int32_t reserved;
    int64_t n_pins;
    const double* x0;       /* [n_mov] initial positions (init_positions_from_bels) */
    const double* y0;       /* [n_mov] */
    const float* area;      /* [n_mov] or NULL */
    const double* fixed_x;  /* [n_fixed] */
    const double* fixed_y;  /* [n_fixed] */
    const uint32_t* net_off;   /* [n_nets + 1] */
    const uint32_t* pin_cell;  /* [n_pins] */
    const double* pinw;        /* [n_mov] or NULL */
    const double* net_weight;  /* [n_nets] or NULL (every net weighs 1): the optional net
                                  weights of add_wa_wirelength_gradient (common.rs:204-261),
                                  e.g. a timing-driven caller's 1 + timing_weight * crit^2;
                                  ElectroPlace passes none (mod.rs:112-115) */
} ep_design;

/* ElectroPlaceCfg (electro_place/config.rs:5-40) plus the constants of
 * mod.rs:31-43 and this library's knobs.  ep_config_default() gives the
 * reference defaults. */
typedef struct ep_config {
    uint64_t seed;              /* 1 */
    double wl_coeff;            /* 0.5 */
    double target_util;         /* 0.7 (unused by the reference loop) */
    int32_t timing_driven;      /* 0 (unused by the reference loop) */
    int32_t max_iters;          /* 500 */
    double target_density;      /* 1.0 (only moves the zeroed DC term) */
    double timing_weight;       /* 0.0 (unused) */
    double nesterov_step_size;  /* 0.1 */
    int32_t legalize_interval;  /* 5 */
    int32_t min_iters;          /* 20  (CONVERGENCE_MIN_ITERS) */
    double overflow_target;     /* 0.1 (OVERLAP_CONVERGE) */
    double density_eta;         /* 0.1 (DENSITY_ETA) */
    double grow_mult;           /* 1.025 */
    double grow_threshold;      /* 50.0 */
    double grow_add;            /* 1.0 */
    int32_t stop_every_iter;    /* 0: reference rule, check on legalize iterations */
    int32_t log_hpwl;           /* 1: continuous HPWL into the log on check iterations */
    int32_t check_nonfinite;    /* 1: fail with EP_ERR_NONFINITE on NaN/Inf */
    int32_t field_mode;         /* EP_FIELD_REFERENCE or EP_FIELD_IDXST */
} ep_config;

/* Per-iteration record (mod.rs:185-188 logs the same quantities). */
typedef struct ep_iter_rec {
    int32_t iter;
    int32_t checked;        /* convergence-check iteration */
    double overflow;        /* compute_overlap at the look-ahead point */
    double lambda;          /* density penalty used this iteration */
    double step_size_x, step_size_y;
    double step_norm_x, step_norm_y;
    double hpwl;            /* continuous HPWL after the step (check iterations) */
    double wl_norm, den_norm; /* norms used by the penalty initialisation */
} ep_iter_rec;

typedef struct ep_result {
    int32_t status;
    int32_t iters;          /* iterations executed */
    int32_t converged;      /* overflow < target on a check iteration */
    int32_t precision;
    double overflow;        /* overflow of the last executed iteration */
    double hpwl;            /* continuous HPWL of the final positions (FP64) */
    double lambda;
    double t_submit_ms;     /* serving: submit, admit and finish times */
    double t_admit_ms;
    double t_done_ms;       /* batch placement: time from the placement's reset
                               to the iteration that stopped the design, on the
                               device clock (submit and admit are 0) */
} ep_result;

/* ---- context ---- */
int ep_create(ep_ctx** out, int device, uint64_t hbm_budget_bytes);
void ep_destroy(ep_ctx* ctx);
const char* ep_last_error(ep_ctx* ctx);
void ep_config_default(ep_config* cfg);
int ep_version(void);

/* ---- resident designs ---- */
int ep_design_upload(ep_ctx* ctx, const ep_design* d, int32_t precision, uint32_t* design_id);
int ep_design_free(ep_ctx* ctx, uint32_t design_id);
uint64_t ep_design_bytes(const ep_design* d, int32_t precision);

/* Runs place_electro's continuous loop (mod.rs:104-197) on one resident
 * design, synchronously. */
int ep_place(ep_ctx* ctx, uint32_t design_id, const ep_config* cfg, ep_result* res);
/* Advances every listed design in the same launches until each stops. */
int ep_place_batch(ep_ctx* ctx, const uint32_t* ids, uint32_t n, const ep_config* cfg,
                   ep_result* results);
/* No counterpart in the reference.  ep_place for callers that cannot
 * allocate device memory (serving loops kept off the allocator, harnesses
 * that forbid it): the same placement, bit for bit, with no hipMalloc, no
 * hipFree and no graph instantiation.  The batch's launch descriptors go
 * into one file-scope device buffer of 1 MiB per GPU (held until the process
 * ends), enough for designs of tens of millions of pins; a design whose
 * descriptors do not fit is refused with EP_ERR_INVALID_ARG.  The design
 * must have been uploaded with ep_design_upload (which prepares its
 * transform matrices).  Calls that overlap on one GPU from different
 * contexts race for the buffer; calls through one context are serialised. */
int ep_place_noalloc(ep_ctx* ctx, uint32_t design_id, const ep_config* cfg, ep_result* res);
int ep_positions(ep_ctx* ctx, uint32_t design_id, double* x, double* y);
int ep_iter_log(ep_ctx* ctx, uint32_t design_id, ep_iter_rec* recs, uint32_t cap, uint32_t* n);

/* ---- single-kernel evaluation on a resident design (host buffers) ---- */
/* Configuration used by the ep_eval_* calls (wl_coeff, field_mode); the
 * reference defaults until set. */
int ep_eval_config(ep_ctx* ctx, const ep_config* cfg);
/* WA gradient (common.rs:205-265, wa.rs:51-101) at positions (x, y). */
int ep_eval_wa_grad(ep_ctx* ctx, uint32_t id, const double* x, const double* y,
                    double* gx, double* gy);
/* Density (density.rs:54-74) and overflow (density.rs:79-86). */
int ep_eval_density(ep_ctx* ctx, uint32_t id, const double* x, const double* y,
                    double* dens, double* overflow);
/* Integer bin ranges {bx_lo, bx_hi, by_lo, by_hi} per cell (density.rs:20-37). */
int ep_eval_bin_ranges(ep_ctx* ctx, uint32_t id, const double* x, const double* y,
                       int32_t* ranges);
/* Field (density.rs:96-166) of a given density grid [By*Bx]. */
int ep_eval_field(ep_ctx* ctx, uint32_t id, const double* dens, double* ex, double* ey);
/* Density gradient gather (density.rs:171-206). */
int ep_eval_density_grad(ep_ctx* ctx, uint32_t id, const double* x, const double* y,
                         const double* ex, const double* ey, double* gx, double* gy);
/* Continuous HPWL (FP64 reduction). */
int ep_eval_hpwl(ep_ctx* ctx, uint32_t id, const double* x, const double* y, double* hpwl);

/* ---- serving: continuous batching under an HBM budget ---- */
typedef struct ep_serve_config {
    uint32_t max_slots;         /* designs advanced per fused iteration (<= 128) */
    uint32_t iters_per_sync;    /* iterations between host polls (5) */
    uint64_t hbm_budget_bytes;  /* 0: the context budget */
} ep_serve_config;

int ep_serve_start(ep_ctx* ctx, const ep_serve_config* scfg);
int ep_serve_submit(ep_ctx* ctx, const ep_design* d, const ep_config* cfg, uint64_t* ticket);
/* 0 (EP_OK) when done, EP_PENDING while queued or running, < 0 on error. */
int ep_serve_poll(ep_ctx* ctx, uint64_t ticket, ep_result* res);
int ep_serve_positions(ep_ctx* ctx, uint64_t ticket, double* x, double* y);
int ep_serve_release(ep_ctx* ctx, uint64_t ticket);
int ep_serve_cancel(ep_ctx* ctx, uint64_t ticket);
/* Counters of the serving thread (no counterpart in the reference).  After a
 * failure inside the serving thread (a device error), every queued and
 * running request ends with that status, `failed` holds it and later
 * ep_serve_submit calls return it. */
typedef struct ep_serve_stats {
    uint64_t admitted;        /* requests admitted into the running batch */
    uint64_t finished;        /* requests that left it (done or cancelled) */
    uint64_t batch_builds;    /* batch rebuilds after membership changes */
    uint64_t graph_captures;  /* iteration graphs captured and instantiated */
    uint64_t graph_updates;   /* instantiated graphs updated in place after a rebuild */
    uint64_t chunks;          /* chunks of iterations launched */
    int32_t failed;           /* 0, or the status that stopped the serving thread */
    int32_t reserved;
} ep_serve_stats;
int ep_serve_stats_get(ep_ctx* ctx, ep_serve_stats* out);
int ep_serve_stop(ep_ctx* ctx);

/* ---- measurement ---- */
typedef struct ep_phase_time {
    char name[24];
    int32_t launches;       /* iterations averaged */
    int32_t reserved;
    double avg_ms;          /* HIP-event time of the phase per iteration */
} ep_phase_time;

/* Resets the listed designs, runs `warm` iterations, then times each phase
 * of `reps` further iterations with HIP events on the context's stream. */
int ep_profile_iteration(ep_ctx* ctx, const uint32_t* ids, uint32_t n, const ep_config* cfg,
                         uint32_t warm, uint32_t reps, ep_phase_time* out, uint32_t cap,
                         uint32_t* n_out);

/* Autotuner (no counterpart in the reference): for each distinct bin shape
 * among the FAST designs `ids`, largest first, times every eligible
 * transform engine (engine 0: MFMA with LDS staging in K stages of 32;
 * 1: FP32 FMA; 2: MFMA split-K inside a block; 3: MFMA with K stages of 64
 * through one LDS buffer;
 * 4: engine 0 with products of K >= 512 split across blocks into K slices
 * of 128; 5: FP32 FFTs of the rows, bin counts that are powers of two from
 * 16 to 512) on the four
 * transform passes of the whole batch `ids`, `reps` times, with the other
 * shapes at their current choice, and keeps the fastest per shape for every
 * later batch.  With a cache path, shapes found in the file are reused and
 * new results appended. */
int ep_autotune(ep_ctx* ctx, const uint32_t* ids, uint32_t n, uint32_t reps, const char* cache_path);
/* The tuning table as (Bx, By, engine) triples; *n receives the entry count. */
int ep_tune_table(ep_ctx* ctx, int32_t* out, uint32_t cap, uint32_t* n);
/* Pins the transform engine of one bin shape in the tuning table (engine < 0
 * removes the entry; the MFMA engines 0, 2, 3, 4 need bins that are multiples
 * of 64, the FFT engine 5 powers of two from 16 to 512).  No counterpart in
 * the reference. */
int ep_tune_set(ep_ctx* ctx, int32_t bx, int32_t by, int32_t engine);

/* Copies `bytes` device-to-device once with the runtime's plain copy kernel
 * (no counterpart in the reference): a transfer of known size against which
 * hardware traffic counters are calibrated. */
int ep_calibrate_copy(ep_ctx* ctx, uint64_t bytes);

/* ---- host helpers ---- */
/* Greedy legalization, mod.rs:206-256: cells in index order take the free
 * BEL of their bucket nearest in squared distance, first minimum in BEL
 * order.  bel_taken marks BELs already bound on entry and is updated. */
int ep_legalize(int32_t n_cells, const double* x, const double* y, const int32_t* cell_bucket,
                int32_t n_bels, const int32_t* bel_x, const int32_t* bel_y,
                const int32_t* bel_bucket, uint8_t* bel_taken, int32_t* cell_bel,
                double* displacement);

/* The legalization view of a design: what `legalize_electro` (mod.rs:206-256)
 * binds and what `total_hpwl` (metrics/wirelength.rs:10-51) measures.
 * Buckets are the cell types; a bucket's BELs in ascending index order are
 * its bucket order (context/buckets.rs:25-43).  Cluster children
 * (common.rs:319-374) are the non-root cells of movable clusters: the solver
 * does not move them, legalization places each right after its root, at
 * root location + (dx, dy) when a free BEL of its bucket is there, else on
 * the first free BEL of its bucket.  A child with pins is one of the
 * design's fixed cells (child_fixed; -1 without pins), whose location the
 * WA gradient reads, so it follows the child after every legalization. */
typedef struct ep_legal_problem {
    int32_t n_bels;
    int32_t n_children;
    const int32_t* bel_x;         /* [n_bels] */
    const int32_t* bel_y;         /* [n_bels] */
    const int32_t* bel_bucket;    /* [n_bels] */
    const uint8_t* bel_taken;     /* [n_bels] or NULL: BELs bound by cells the placer never
                                     unbinds (locked cells) */
    const int32_t* cell_bucket;   /* [n_mov] bucket of each movable cell */
    const int32_t* child_root;    /* [n_children] solver index of the root; roots ascending,
                                     a root's children in their constraint order */
    const int32_t* child_dx;      /* [n_children] constr_x */
    const int32_t* child_dy;      /* [n_children] constr_y */
    const int32_t* child_bucket;  /* [n_children] */
    const int32_t* child_fixed;   /* [n_children] fixed-cell index or -1 */
    const uint8_t* net_has_driver; /* [n_nets] or NULL (every net has one): driverless nets
                                      count 0 in the legalized HPWL (wirelength.rs:16-18) */
} ep_legal_problem;

/* legalize_electro on an ep_legal_problem (mod.rs:206-256 with
 * place_cluster_children, common.rs:319-374); child_bel may be NULL when
 * n_children is 0.  EP_ERR_NO_BELS / EP_ERR_PLACEMENT_FAILED as the
 * reference's PlacerError. */
int ep_legalize_problem(const ep_legal_problem* lp, int32_t n_cells, const double* x, const double* y,
                        int32_t* cell_bel, int32_t* child_bel, double* displacement);
/* total_hpwl (metrics/wirelength.rs:10-51) of a legalized binding of d. */
int ep_legal_hpwl(const ep_design* d, const ep_legal_problem* lp, const int32_t* cell_bel,
                  const int32_t* child_bel, double* hpwl);

typedef struct ep_legal_result {
    int32_t best_iter;         /* iteration of the best snapshot; -1: the initial positions */
    int32_t legalizations;     /* snapshots legalized and scored */
    double best_hpwl;          /* its legalized HPWL (+inf when none was scored) */
    double final_displacement; /* of legalize(best) */
} ep_legal_result;

/* place_electro (mod.rs:65-204) end to end on a resident design: the
 * continuous loop on the device, and at every check iteration (iter %
 * legalize_interval == 0, the last iteration, and the iteration a design
 * stops at) the positions legalized on the host, scored by total_hpwl and
 * kept when strictly better (NesterovLoopState::record_metric,
 * common.rs:485-495); cluster children with pins move to their new BELs
 * before the next iteration.  The final binding is legalize(best)
 * (mod.rs:199) into cell_bel [n_mov] and child_bel [n_children]. */
int ep_place_legalized(ep_ctx* ctx, uint32_t design_id, const ep_config* cfg,
                       const ep_legal_problem* lp, ep_result* res, ep_legal_result* lres,
                       int32_t* cell_bel, int32_t* child_bel);

/* Synthetic generic-FPGA design generator (clustered-locality netlists). */
typedef struct ep_gen_params {
    uint64_t seed;
    int32_t n_cells;        /* movable */
    int32_t n_fixed;        /* IO cells spaced evenly on the perimeter */
    int32_t n_nets;         /* nets with the 60/25/12/3 % degree mix */
    int32_t W, H, Bx, By;
    int32_t n_clusters;
    int32_t area_mix;       /* 0: unit cells; 1: areas 1/2/4 at 85/12/3 % */
    int32_t n_types;        /* movable cell types (BEL buckets), 1 or 2 */
    int32_t n_high_fanout;  /* extra nets of hf_min..hf_max pins */
    int32_t hf_min, hf_max;
    double locality;        /* fraction of a net's pins from its cluster (0.8) */
    double type1_frac;      /* fraction of cells of type 1 (0.2 => 1200 LUT + 300 DFF) */
} ep_gen_params;

typedef struct ep_gen_design {
    ep_design d;            /* arrays below, owned by the generator */
    int32_t n_types;
    int32_t n_bels;         /* logic BELs: one per type per tile, tile-major */
    int32_t* cell_type;     /* [n_mov] */
    int32_t* bel_x;         /* [n_bels] */
    int32_t* bel_y;
    int32_t* bel_type;
    int32_t* init_bel;      /* [n_mov] BEL of the seeded initial placement */
} ep_gen_design;

int ep_generate(const ep_gen_params* p, ep_gen_design** out);
void ep_gen_free(ep_gen_design* g);

#ifdef __cplusplus
}
#endif
#endif
