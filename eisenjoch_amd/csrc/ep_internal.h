// Internal layout shared by the HIP kernels and the host runtime.
//
// One resident design owns one device arena.  Every per-iteration kernel
// covers a whole batch of designs: a block finds its design and its block
// index inside the design in a per-grid block table, and skips the design when its
// state says it is done.  All control scalars are FP64 and live on the
// device, so a batch advances without host round trips.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "eplace.h"

namespace ep {

constexpr int kBlock = 256;          // threads per block for cell/bin kernels
constexpr int kWave = 64;            // CDNA wavefront
constexpr int kMaxBuckets = 6;       // group sizes 2,4,8,16,32,64
constexpr uint32_t kPad = 0xFFFFFFFFu;  // empty pin slot
constexpr int kCellParts = 8;        // partial sums per cell block
constexpr int kBinParts = 2;         // partial sums per bin block
constexpr double kFixScale = 2147483648.0;  // 2^31 fixed point for deterministic sums (contributions <= 1)
constexpr int kScatCells = 1024;     // cells per scatter block
constexpr int kScatBins = 6144;      // LDS accumulation window of a scatter block (bins)
constexpr int kHfStat = 16;      // per high-fanout net: 4 extremes, 8 sums, weight, pad
constexpr int kXcds = 8;          // MI355X: blocks are dealt to the 8 XCDs in turn
constexpr int kRefBits = 13;      // design bits of a block-table entry (entry = design | block << 13)
constexpr int kCellSlots = 12;    // slots of a cell held at a fixed stride (cs8)
constexpr int kWaUnits = 3;       // 64-slot net units per wave of the WA kernel
// blocks (four waves) of the WA kernel for a design of n_wa_waves units
inline int wa_blocks_of(int n_wa_waves) { return (n_wa_waves + 4 * kWaUnits - 1) / (4 * kWaUnits); }
constexpr int kSplitMinK = 512;   // K from which transform products are split
constexpr int kSplitSlice = 128;  // K per slice of a split product

// Field modes (the transform constants differ, nothing else).
enum FieldMode : int32_t { kFieldReference = 0, kFieldIdxst = 1 };

// Device-resident control state of one design (one per design).
struct DevState {
    double lam;           // density penalty of the current iteration
    double s[2];          // step sizes (x, y)
    double a;             // FISTA a_k
    double m;             // momentum of this iteration's step
    double m_next;        // momentum of the next look-ahead
    double ov;            // overflow of this iteration
    double wl_norm, den_norm;
    double hpwl;          // last continuous HPWL
    double final_hpwl;
    int32_t it;           // iteration being executed (advanced by ctrl_b)
    int32_t lam_init;     // penalty initialised
    int32_t need_fix;     // this iteration's penalty was set by ctrl_a
    int32_t done;         // skip all work
    int32_t run_update;   // the update of this iteration runs
    int32_t check;        // this iteration is a convergence-check iteration
    int32_t converged;
    int32_t nonfinite;
    int32_t prev_pending; // step-norm / hpwl partials of the last update await reduction
    int32_t prev_check;
    int32_t n_log;
    int32_t pad0;
    uint64_t t_start;     // device wall clock at the placement's reset
    uint64_t t_done;      // device wall clock at the control step that stopped the design
};

// Constant per-design configuration (copied into the descriptor).
struct DevCfg {
    double wl_coeff, overflow_target, eta, grow_mult, grow_thr, grow_add, step0;
    int32_t max_iters, min_iters, interval, stop_every_iter, log_hpwl, check_nonfinite;
};

// Descriptor of one resident design; pointer fields are typed by precision
// (T = double for EP_PREC_FP64, float for EP_PREC_FAST).
struct DesignDev {
    int32_t n_mov, n_fixed, n_nets, W, H, Bx, By, prec;
    int32_t n_cell_blocks, n_bin_blocks, n_wa_waves, n_hf, n_buckets, n_scat_blocks;
    int32_t n_slots;              // per-pin gradient slots: n_wa_waves * 64 + high-degree pins
    int32_t n_keys;               // Morton keys of the 8x8-bin tiles (a power of four)
    int32_t tiles_x, tiles_y;     // 8x8-bin tiles covering the bin grid
    int32_t scat_pull;            // 1: tile-owned pull scatter (footprints span <= 9 bins)
    int32_t n_tile_blocks;        // pull scatter blocks (tiles_x * tiles_y, or 0)
    int32_t n_ov_parts;           // overflow partial pairs written per iteration
    int32_t bucket_wave_start[kMaxBuckets + 1];
    int32_t bucket_group[kMaxBuckets];
    double rx, ry, max_x, max_y;
    DevCfg cfg;
    DevState* st;
    ep_iter_rec* log;          // [max_iters]
    // cells
    void* x0[2];               // initial positions (reset source)
    void* pos[2];              // x, y: the solver state [n_mov], then the fixed cells' locations
                               // [n_fixed], so a pin of any cell reads pos[ax][cell]
    void* dpos[2];             // x - x_prev
    void* look;                // T [n_mov + n_fixed][2]: clamped look-ahead, x and y interleaved,
                               // then the fixed cells' locations
    void* g[2][2];             // [parity][axis] combined gradient (or density grad)
    void* pg;                  // per-slot WA gradient (gx, gy) in T [n_slots + 1][2]; entry n_slots
                               // is never written and stays zero
    uint32_t* cs_off;          // cell -> slot CSR [n_mov + 1]
    void* gwl;                 // T [n_mov][2] WA gradient per cell (its slots summed in FP64, stored in T)
    const void* slot_w;        // T [n_slots]: each slot's net weight, or null (all nets weigh 1)
    uint32_t* cs_idx;          // slots of each movable cell, ascending (nets of <= 64 pins)
    uint32_t* cs8;             // [n_mov][kCellSlots]: the first kCellSlots of them, then slot n_slots
    uint32_t* chf_off;         // cell -> high-fanout-net CSR [n_mov + 1]
    uint32_t* chf_net;         // high-fanout nets of each cell, ascending
    void* hf_stat;             // T [n_hf][kHfStat]: extremes, exponential sums, weight
    uint32_t* perm[2];         // cells in Morton order of their lower-left bin tile; the
                               // sort of iteration k reads perm[k&1] and writes perm[(k+1)&1]
    uint32_t* perm0;           // [n_mov] the order a reset starts from: the cells by the tile of
                               // their initial position, so the first sort's waves see few keys
    uint32_t* keys;            // tile key of perm[k&1][j], by position j
    void* stage_xy;            // T [n_mov][2]: look-ahead position of perm[k&1][j], by j
    float* stage_hs;           // half side of perm[k&1][j], by j
    void* sorted_xy;           // T [n_mov][2]: look-ahead positions in the new sort order
    float* sorted_hs[2];       // half sides in the order of perm[0] / perm[1]
    uint32_t* tcount;          // [n_keys] cells per tile key (zero between sorts)
    uint32_t* tstart;          // [n_keys + 1] first position of each key in perm
    uint32_t* tofs;            // [n_mov] offset of perm[k&1][j] within its tile, by j
    void* pinw;                // T[n_mov]
    float* hs;                 // half side of the footprint [n_mov]
    // nets
    uint32_t* slots;           // padded pin slots, n_wa_waves * 64
    uint32_t* hf_off;          // [n_hf + 1]  (nets with > 64 pins)
    uint32_t* hf_pins;
    // bins
    void* dacc;                // wide-cell scatter accumulator: FAST u64 fixed point 2^-31 [By*Bx];
                               // FP64 exact 128-bit fixed point, integer words [By*Bx] then
                               // 2^-64 fraction words [By*Bx]
    void* dens;                // T [By*Bx]
    // transform intermediates; a pass whose products are split along K
    // writes ks planes of partial sums (ks_x for K = Bx, ks_y for K = By)
    void* tT;                  // F1 output, transposed [ks_x][Bx*By]
    void* ehat[2];             // F2 output [ks_y][By*Bx]
    void* uT[2];               // I1 output, transposed [ks_x][Bx*By]
    void* field;               // I2 output, T [ks_y][By*Bx][2]: Ex and Ey interleaved; plane 0
                               // holds the sum after the transforms
    int32_t ks_x, ks_y;        // K slices of the products with K = Bx / K = By (1: unsplit)
    void* spec_x;              // T[2*Bx]: (w, 2(1-cos w)) per kx
    void* spec_y;              // T[2*By]
    // reductions
    double* part_cell;         // [n_cell_blocks * kCellParts]
    double* part_bin;          // [n_bin_blocks * kBinParts]
    double* part_net;          // [n_wa_waves + n_hf]
};

// One GEMM of a transform pass: C = A (M x K, row-major) * Bt^T (Bt is
// N x K row-major).  Operand kinds: 0 = data in T (split to BF16 hi/lo on
// the fly by the MFMA engine), 1 = constant (T copy plus BF16 hi/lo planes).
struct GemmJob {
    const void* A;             // T-typed operand (data, or the constant in T)
    const void* A_hi;          // FAST constants: bf16 hi and lo planes
    const void* A_lo;
    const void* B;
    const void* B_hi;
    const void* B_lo;
    void* C;
    void* C2;
    int32_t M, N, K;
    int32_t a_kind, b_kind;
    int32_t epi;               // 0: store, 1: store transposed, 2: Poisson epilogue
    int32_t design;            // index into the descriptor array
    int32_t tiles_n;           // tiles along N
    int32_t tile_start;        // first global block of this job
    int32_t ksplit;            // K slices (one block per tile and slice), MFMA engines 0/3
    int32_t a_parts, b_parts;  // partial planes summed, in order, when loading a data operand
    int32_t pad;
    const DevState* st;        // the design's state (a block of a done design returns at once)
    int64_t c_pstride;         // elements between the output's partial planes (C and C2)
    int64_t a_pstride, b_pstride;  // elements between a data operand's planes
};

// Sum of the partial planes of a split pass, in place into plane 0.
struct PlaneSum {
    float* p;
    int64_t n;        // elements per plane
    int32_t planes;
    int32_t blocks;   // blocks of 256 threads covering n
};

// One batch of row transforms of the FFT engine (FAST path, power-of-two
// row lengths): rows x n input, row-major.  kind: 0 DCT-II, 1 DCT-III with
// the k = 0 term halved (C3), 2 DST-III of the reference field (S3), 3 the
// IDXST sine transform.  Row passes: epi 0 stores element k of row r at
// out[(k * ld + r) * ostride] (the plane transposed), epi 2 at
// out[(r * n + k) * ostride] (row order; ostride 2 writes one component of
// an interleaved pair).  Column-solve jobs (the second pass): DCT-II, the
// Poisson solve, then C3 into out and `kind` into out2, both transposed.
// Pair jobs (the last pass): rows of in through `kind` and rows of in2
// through C3, stored in row order as pairs into out.  A block transforms
// rpb rows.
struct FftJob {
    const void* in;
    const void* in2;     // the last pass: the second component's plane
    void* out;
    void* out2;
    const void* tw;      // float2[n]: exp(-2 pi i t k1 / n) at k1 N2 + t for the pass's split n = N1 N2
    const void* dtw;     // float2[n]: exp(-i pi k / 2n)
    const DevState* st;  // a block of a done design returns at once
    int32_t n, rows, rpb, kind, epi, ld, ostride, design;
};

// kEpiStore2: row-major into every other element (one component of a pair)
enum Epilogue : int32_t { kEpiStore = 0, kEpiStoreT = 1, kEpiPoisson = 2, kEpiStore2 = 3 };

// A batch as the kernels see it (device pointers).
struct BatchDev {
    const DesignDev* d;
    int32_t nd;
    const int32_t* pfx_cell;   // [nd+1] block prefix for cell kernels
    const int32_t* pfx_bin;
    const int32_t* pfx_wa;     // blocks of 4 waves
    const int32_t* pfx_hf;     // one block per high-degree net
    const int32_t* pfx_scat;   // atomic scatter blocks (kScatCells cells each)
    const int32_t* pfx_tile;   // pull scatter blocks (one per 8x8-bin tile)
    // each block's (design, block of the design) in the same six grids, packed
    // as design | block << kRefBits, or -1 for a padding block (one load at
    // block start; null in a one-design view, where the prefix gives both)
    const int32_t* blk_cell;
    const int32_t* blk_bin;
    const int32_t* blk_wa;
    const int32_t* blk_hf;
    const int32_t* blk_scat;
    const int32_t* blk_tile;
};

// ---- launchers (ep_kernels.hip) ----
void launch_wa(const BatchDev& b, int prec, int total_wa_blocks, int total_hf_blocks,
               hipStream_t s);
void launch_scatter(const BatchDev& b, int prec, int total_scat_blocks, hipStream_t s);
void launch_dens_final(const BatchDev& b, int prec, int total_bin_blocks, hipStream_t s);
void launch_gather(const BatchDev& b, int prec, int total_cell_blocks, int mode, hipStream_t s);
void launch_wl_cell(const BatchDev& b, int prec, int total_cell_blocks, hipStream_t s);
void launch_ctrl(const BatchDev& b, int which, hipStream_t s);
void launch_update(const BatchDev& b, int prec, int total_cell_blocks, hipStream_t s);
void launch_hpwl(const BatchDev& b, int prec, int total_wa_blocks, int total_hf_blocks,
                 int final_pass, hipStream_t s);
void launch_reset(const BatchDev& b, int prec, int total_cell_blocks, hipStream_t s);
void launch_tile_sort(const BatchDev& b, int prec, int total_cell_blocks, hipStream_t s);
void launch_scatter_pull(const BatchDev& b, int prec, int total_tile_blocks, hipStream_t s);
void launch_bin_ranges(const BatchDev& b, int total_cell_blocks, int32_t* out, hipStream_t s);
void launch_gather_export(const BatchDev& b, int prec, int total_cell_blocks, double* gx,
                          double* gy, hipStream_t s);
void launch_wa_reduce(const BatchDev& b, int prec, int total_cell_blocks, double* gx, double* gy,
                      hipStream_t s);

// ---- transforms (ep_gemm.hip) ----
// Runs one pass (0 F1, 1 F2, 2 I1, 3 I2; all jobs) of the spectral solve
// on one engine (0 MFMA, 1 FMA, 2 MFMA split-K, 3 MFMA K stages of 64, 4 K slices).
void launch_plane_sum(const PlaneSum* sums, int n_sums, int max_blocks, hipStream_t s);
// bjob: (job, tile) of each block of the grid (total_tiles pairs; -1: none).
void launch_gemm_pass(const GemmJob* jobs_dev, const int32_t* bjob, int total_tiles,
                      const DesignDev* dd, int prec, int engine, int pass, hipStream_t s);

// ---- transforms by FFT (ep_fft.hip) ----
// Row lengths the FFT engine takes: powers of two from 16 to 512.
inline bool fft_len_ok(int n) { return n >= 16 && n <= 512 && (n & (n - 1)) == 0; }
// Rows one block of the FFT pass transforms (256 threads, 4 to 32 threads
// per row).
int fft_rows_per_block(int n);
// fblk: (job, first row) of each block of the grid (blocks pairs; -1: none).
void launch_fft_pass(const FftJob* jobs_dev, const int32_t* fblk, int blocks, const DesignDev* dd,
                     hipStream_t s);
// The column solve (the FFT engine's second pass); rows per block of it.
void launch_fft_solve(const FftJob* jobs_dev, const int32_t* fblk, int blocks, const DesignDev* dd,
                      hipStream_t s);
int fft_solve_rows_per_block(int n);
// The twiddles e^{-2 pi i t k1 / n} of the row passes' split of n (or the
// column solve's) into out (float2[n]), at k1 N2 + t.
void fft_twiddle_table(int n, bool solve, float* out);
// The last pass, both components per block (pair jobs).
void launch_fft_pair(const FftJob* jobs_dev, const int32_t* fblk, int blocks, hipStream_t s);

}  // namespace ep
