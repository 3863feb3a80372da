// Batches, the placement loop (mod.rs:104-197) and single-kernel evaluation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "ep_internal.h"
#include "ep_runtime.h"

namespace ep {

#define HIPCHK(ctx, expr)                                                             \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            return fail(ctx, EP_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

static size_t align_up(size_t v) { return (v + 255) & ~(size_t)255; }

static DevCfg make_cfg(const ep_config* c) {
    DevCfg d;
    d.wl_coeff = c->wl_coeff;
    d.overflow_target = c->overflow_target;
    d.eta = c->density_eta;
    d.grow_mult = c->grow_mult;
    d.grow_thr = c->grow_threshold;
    d.grow_add = c->grow_add;
    d.step0 = c->nesterov_step_size;
    d.max_iters = c->max_iters;
    d.min_iters = c->min_iters;
    d.interval = c->legalize_interval > 0 ? c->legalize_interval : 1;
    d.stop_every_iter = c->stop_every_iter;
    d.log_hpwl = c->log_hpwl;
    d.check_nonfinite = c->check_nonfinite;
    return d;
}

int validate_config(Ctx* c, const ep_config* cfg, int log_cap) {
    if (!cfg) return fail(c, EP_ERR_INVALID_ARG, "null config");
    if (cfg->max_iters < 0 || cfg->max_iters > log_cap)
        return fail(c, EP_ERR_INVALID_ARG, "max_iters out of range (0.." + std::to_string(log_cap) + ")");
    if (!(cfg->wl_coeff > 0.0) || !std::isfinite(cfg->wl_coeff))
        return fail(c, EP_ERR_INVALID_ARG, "wl_coeff must be positive");
    if (cfg->field_mode != EP_FIELD_REFERENCE && cfg->field_mode != EP_FIELD_IDXST)
        return fail(c, EP_ERR_INVALID_ARG, "unknown field mode");
    return EP_OK;
}

// Transform engine of a design shape: the autotuned (or pinned) choice when
// there is one, else the FFT engine for bin counts that are powers of two up
// to 512, MFMA with LDS staging in K stages of 32 for other shapes whose bins
// are multiples of 64, and the FMA kernel otherwise.  The FP64 path always
// takes the FMA kernel.
int engine_for(Ctx* c, int prec, int Bx, int By) {
    if (prec != EP_PREC_FAST || getenv("EP_FORCE_FMA") != nullptr) return 1;
    auto it = c->tune.find({Bx, By});
    if (it != c->tune.end()) return it->second;
    if (fft_len_ok(Bx) && fft_len_ok(By)) return kEngineFft;
    return (Bx % 64 == 0 && By % 64 == 0) ? 0 : 1;
}

// Engines a FAST design shape can take.
static std::vector<int> engines_for_shape(int Bx, int By) {
    std::vector<int> e;
    if (Bx % 64 == 0 && By % 64 == 0) e = {0, 3, 4, 2};
    e.push_back(1);
    if (fft_len_ok(Bx) && fft_len_ok(By)) e.push_back(kEngineFft);
    return e;
}

// Block table of one grid (P: the per-design block prefix).  Entry p is
// design | (block << kRefBits), or -1 for a block that pads the XCDs to
// equal length.  The dispatcher deals consecutive blocks to the kXcds XCDs
// in turn, so position p runs on XCD p % kXcds: a design with xcd_of[d] >= 0
// has all its blocks at that XCD's positions (its data is then fetched into
// one L2), a design with xcd_of[d] < 0 has its blocks dealt over every XCD.
static void build_block_table(const std::vector<int32_t>& P, int nd, const std::vector<int>& xcd_of,
                              std::vector<int32_t>& tab) {
    std::vector<std::vector<int32_t>> lists(kXcds);
    int turn = 0;
    for (int d = 0; d < nd; d++) {
        const int n = P[d + 1] - P[d];
        for (int j = 0; j < n; j++) {
            const int x = d < (int)xcd_of.size() && xcd_of[d] >= 0 ? xcd_of[d] % kXcds : (turn++ % kXcds);
            lists[x].push_back((int32_t)d | ((int32_t)j << kRefBits));
        }
    }
    size_t len = 0;
    for (auto& l : lists) len = std::max(len, l.size());
    tab.assign(std::max<size_t>(1, len * kXcds), -1);
    for (int x = 0; x < kXcds; x++)
        for (size_t q = 0; q < lists[x].size(); q++) tab[q * kXcds + x] = lists[x][q];
    if (P[nd] == 0) tab.assign(1, -1);
}

// XCD of each design: designs whose work (cells plus pins) is at most an
// XCD's share of the batch go whole to the least loaded XCD, largest first,
// so the random reads of their kernels (pin positions, slot values, the
// tile count's positions) are served from that XCD's L2; larger designs are
// dealt over every XCD.  The FFT passes keep every design on every XCD:
// their work follows the bin count, which this weight does not balance.
static std::vector<int> assign_xcds(const std::vector<Design*>& ds) {
    const int nd = (int)ds.size();
    std::vector<int> xcd(nd, -1);
    std::vector<double> w(nd, 0.0);
    double tot = 0.0;
    for (int d = 0; d < nd; d++) {
        w[d] = (double)ds[d]->dev.n_mov + (double)(ds[d]->cs_count + ds[d]->hf_pin_count);
        tot += w[d];
    }
    std::vector<int> order(nd);
    for (int d = 0; d < nd; d++) order[d] = d;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b2) { return w[a] > w[b2]; });
    std::vector<double> load(kXcds, 0.0);
    for (int d : order) {
        if (w[d] > tot / kXcds) {
            for (double& l : load) l += w[d] / kXcds;
            continue;
        }
        const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        load[x] += w[d];
        xcd[d] = x;
    }
    return xcd;
}

// Per pass: (A kind, A matrix or buffer, B kind, B matrix or buffer, epilogue, out, out2).
// The device image (descriptors, prefixes, jobs, block tables) goes into a
// fresh allocation, or, with ext_mem, into that caller-owned device buffer of
// ext_cap bytes (EP_ERR_INVALID_ARG when it does not fit; nothing is allocated).
int batch_build_cfgs(Ctx* c, const std::vector<Design*>& ds, const std::vector<const ep_config*>& cfgs,
                     Batch& b, void* ext_mem, size_t ext_cap) {
    // a batch that keeps its graphs hands them to its next membership, stale
    hipGraphExec_t kept[2] = {nullptr, nullptr};
    int kept_iters[2] = {0, 0};
    if (b.keep_graphs)
        for (int k = 0; k < 2; k++) {
            kept[k] = b.graph[k];
            kept_iters[k] = b.graph_iters[k];
            b.graph[k] = nullptr;
        }
    batch_release(b);
    for (int k = 0; k < 2; k++) {
        b.graph[k] = kept[k];
        b.graph_iters[k] = kept_iters[k];
        b.graph_stale[k] = kept[k] != nullptr;
    }
    b.ds = ds;
    const int nd = (int)ds.size();
    if (nd == 0) return EP_OK;
    b.prec = ds[0]->prec;
    b.log_hpwl = 0;
    for (int i = 0; i < nd; i++) {
        if (ds[i]->prec != b.prec) return fail(c, EP_ERR_INVALID_ARG, "mixed precisions in one batch");
        if (cfgs[i]->log_hpwl) b.log_hpwl = 1;
    }
    std::vector<DesignDev> dv(nd);
    std::vector<int32_t> pc(nd + 1), pb(nd + 1), pw(nd + 1), ph(nd + 1), ps(nd + 1), pt(nd + 1);
    pc[0] = pb[0] = pw[0] = ph[0] = ps[0] = pt[0] = 0;
    // jobs: [pass][engine]
    std::vector<GemmJob> jobs[4][kEngines];
    std::vector<FftJob> fjobs[4];  // the FFT engine's row transforms, per pass
    std::vector<PlaneSum> sums;  // split final passes: planes summed after the transforms
    for (int i = 0; i < nd; i++) {
        Design& D = *ds[i];
        dv[i] = D.dev;
        dv[i].cfg = make_cfg(cfgs[i]);
        pc[i + 1] = pc[i] + D.dev.n_cell_blocks;
        pb[i + 1] = pb[i] + D.dev.n_bin_blocks;
        pw[i + 1] = pw[i] + wa_blocks_of(D.dev.n_wa_waves);
        ph[i + 1] = ph[i] + D.dev.n_hf;
        ps[i + 1] = ps[i] + D.dev.n_scat_blocks;
        pt[i + 1] = pt[i] + D.dev.n_tile_blocks;
        Transform *tx, *ty;
        int rc = get_transform(c, D.dev.Bx, cfgs[i]->field_mode, &tx);
        if (!rc) rc = get_transform(c, D.dev.By, cfgs[i]->field_mode, &ty);
        if (rc) return rc;
        const bool f64 = b.prec == EP_PREC_FP64;
        const size_t tsz = f64 ? 8 : 4;  // element size of the data buffers
        auto cst = [&](GemmJob& J, bool isA, Transform* t, int k) {
            const void* p = f64 ? (const void*)t->d[k] : (const void*)t->f[k];
            if (isA) { J.A = p; J.A_hi = t->hi[k]; J.A_lo = t->lo[k]; J.a_kind = 1; }
            else { J.B = p; J.B_hi = t->hi[k]; J.B_lo = t->lo[k]; J.b_kind = 1; }
        };
        auto dat = [&](GemmJob& J, bool isA, const void* p) {
            if (isA) { J.A = p; J.A_hi = J.A_lo = nullptr; J.a_kind = 0; }
            else { J.B = p; J.B_hi = J.B_lo = nullptr; J.b_kind = 0; }
        };
        const int Bx = D.dev.Bx, By = D.dev.By;
        auto mk = [&](int M, int N, int K, int epi, void* C, void* C2) {
            GemmJob J;
            memset(&J, 0, sizeof(J));
            J.M = M; J.N = N; J.K = K; J.epi = epi; J.C = C; J.C2 = C2; J.design = i;
            J.st = D.dev.st;
            J.tiles_n = (N + 63) / 64;
            J.ksplit = 1;
            J.a_parts = J.b_parts = 1;
            return J;
        };
        GemmJob f1 = mk(By, Bx, Bx, kEpiStoreT, D.dev.tT, nullptr);  // d . C2^T   (rows)
        dat(f1, true, D.dev.dens); cst(f1, false, tx, 0);
        GemmJob f2 = mk(By, Bx, By, kEpiPoisson, D.dev.ehat[0], D.dev.ehat[1]);  // C2 . T (cols)
        cst(f2, true, ty, 0); dat(f2, false, D.dev.tT);
        GemmJob i1a = mk(By, Bx, Bx, kEpiStoreT, D.dev.uT[0], nullptr);  // Ex^ . S3^T (idst_rows)
        dat(i1a, true, D.dev.ehat[0]); cst(i1a, false, tx, 2);
        GemmJob i1b = mk(By, Bx, Bx, kEpiStoreT, D.dev.uT[1], nullptr);  // Ey^ . C3^T (idct_rows)
        dat(i1b, true, D.dev.ehat[1]); cst(i1b, false, tx, 1);
        GemmJob i2a = mk(By, Bx, By, kEpiStore2, D.dev.field, nullptr);  // C3 . U (idct_cols)
        cst(i2a, true, ty, 1); dat(i2a, false, D.dev.uT[0]);
        GemmJob i2b = mk(By, Bx, By, kEpiStore2, (char*)D.dev.field + tsz, nullptr);  // S3 . U (idst_cols)
        cst(i2b, true, ty, 2); dat(i2b, false, D.dev.uT[1]);
        const int eng = engine_for(c, b.prec, Bx, By);
        if (eng == kEngineFft) {
            // three passes over the same planes: rows along x (DCT-II, stored
            // transposed), the column solve along y (DCT-II, Poisson, C3 for
            // Ex and the sine kind for Ey, stored transposed), rows along x
            // (the sine kind for Ex, C3 for Ey) into the field in row order
            const int sk = cfgs[i]->field_mode == EP_FIELD_REFERENCE ? 2 : 3;
            auto fj = [&](const void* in, void* out, void* out2, const Transform* t, int n, int rows,
                          int kind, int epi, int ld, int os, int rpb, bool solve) {
                FftJob F;
                memset(&F, 0, sizeof(F));
                F.in = in; F.out = out; F.out2 = out2;
                F.tw = solve ? t->fft_tw_solve : t->fft_tw; F.dtw = t->dct_tw; F.st = D.dev.st;
                F.n = n; F.rows = rows; F.rpb = std::min(rows, rpb);
                F.kind = kind; F.epi = epi; F.ld = ld; F.ostride = os; F.design = i;
                return F;
            };
            fjobs[0].push_back(fj(D.dev.dens, D.dev.tT, nullptr, tx, Bx, By, 0, 0, By, 1,
                                  fft_rows_per_block(Bx), false));
            fjobs[1].push_back(fj(D.dev.tT, D.dev.uT[0], D.dev.uT[1], ty, By, Bx, sk, 0, Bx, 1,
                                  fft_solve_rows_per_block(By), true));
            FftJob last = fj(D.dev.uT[0], D.dev.field, nullptr, tx, Bx, By, sk, 2, Bx, 2,
                             fft_rows_per_block(Bx), false);
            last.in2 = D.dev.uT[1];
            fjobs[3].push_back(last);
            continue;
        }
        if (eng == 4) {
            // split the long products along K (planes of partial sums); the
            // next pass sums a data operand's planes as it loads it
            const int64_t pl = (int64_t)Bx * By;
            const int sx = D.dev.ks_x, sy = D.dev.ks_y;
            f1.ksplit = sx; f1.c_pstride = pl;
            f2.ksplit = sy; f2.c_pstride = pl; f2.b_parts = sx; f2.b_pstride = pl;
            for (GemmJob* J : {&i1a, &i1b}) {
                J->ksplit = sx; J->c_pstride = pl; J->a_parts = sy; J->a_pstride = pl;
            }
            for (GemmJob* J : {&i2a, &i2b}) {
                J->ksplit = sy; J->c_pstride = 2 * pl; J->b_parts = sx; J->b_pstride = pl;
            }
            if (sy > 1) {
                PlaneSum ps;
                ps.p = (float*)D.dev.field;
                ps.n = 2 * pl;
                ps.planes = sy;
                ps.blocks = (int)((ps.n + 255) / 256);
                sums.push_back(ps);
            }
        }
        jobs[0][eng].push_back(f1);
        jobs[1][eng].push_back(f2);
        jobs[2][eng].push_back(i1a);
        jobs[2][eng].push_back(i1b);
        jobs[3][eng].push_back(i2a);
        jobs[3][eng].push_back(i2b);
    }
    // device image: descriptors, prefixes, jobs
    size_t bytes = align_up(sizeof(DesignDev) * nd) + 6 * align_up(sizeof(int32_t) * (nd + 1)) + align_up(4);
    // Block map of each pass: the dispatcher deals consecutive blocks to the
    // kXcds XCDs in turn, so block b runs on XCD b % kXcds.  The tiles that
    // share a panel of a product's data operand (a row of output tiles when
    // the data is A, in F1 and I1; a column when it is B, in F2 and I2; per K
    // slice) go whole to one XCD, dealt largest first to the least loaded
    // XCD (work = tiles x K stages): the panel is then fetched into one L2
    // and the XCDs stay balanced.  Entries are (job, tile), or (-1, -1) for
    // the blocks that pad the XCDs to equal length.
    std::vector<int32_t> bjob[4][kEngines];
    for (int p = 0; p < 4; p++)
        for (int e = 0; e < kEngines; e++) {
            // a unit: the (job, tile) pairs that share a data panel, and its work
            struct Unit { std::vector<int32_t> jt; int64_t w; };
            std::vector<Unit> units;
            const bool by_col = (p & 1) != 0;
            int t0 = 0;
            std::vector<int> start(jobs[p][e].size());
            for (size_t q = 0; q < jobs[p][e].size(); q++) {
                GemmJob& J = jobs[p][e][q];
                J.tile_start = t0;
                start[q] = t0;
                t0 += ((J.M + 63) / 64) * J.tiles_n * J.ksplit;
            }
            for (size_t q = 0; q < jobs[p][e].size(); q++) {
                const GemmJob& J = jobs[p][e][q];
                const int tm = (J.M + 63) / 64, tiles = tm * J.tiles_n;
                const int64_t stages = std::max(1, J.K / J.ksplit / 32);
                // I2 writes Ex and Ey interleaved from two jobs of one design
                // (pushed one after the other): their tiles of a column form one
                // unit, so the two halves of each output line are written by
                // neighbouring blocks of one XCD and merge in its L2
                const bool pair = p == 3 && J.epi == kEpiStore2 && q + 1 < jobs[p][e].size() &&
                                  jobs[p][e][q + 1].design == J.design && jobs[p][e][q + 1].epi == kEpiStore2;
                for (int ks = 0; ks < J.ksplit; ks++) {
                    if (by_col)
                        for (int n = 0; n < J.tiles_n; n++) {
                            Unit u{{}, 0};
                            for (int m = 0; m < tm; m++)
                                for (size_t h = 0; h < (pair ? 2u : 1u); h++) {
                                    u.jt.push_back((int32_t)(q + h));
                                    u.jt.push_back(start[q + h] + ks * tiles + m * J.tiles_n + n);
                                    u.w += stages;
                                }
                            units.push_back(std::move(u));
                        }
                    else
                        for (int m = 0; m < tm; m++) {
                            Unit u{{}, 0};
                            for (int n = 0; n < J.tiles_n; n++) {
                                u.jt.push_back((int32_t)q);
                                u.jt.push_back(start[q] + ks * tiles + m * J.tiles_n + n);
                                u.w += stages;
                            }
                            units.push_back(std::move(u));
                        }
                }
                if (pair) q++;
            }
            std::stable_sort(units.begin(), units.end(), [](const Unit& a, const Unit& r) { return a.w > r.w; });
            std::vector<std::vector<int32_t>> lists(kXcds);
            std::vector<int64_t> load(kXcds, 0);
            for (const Unit& u : units) {
                const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
                load[x] += u.w;
                lists[x].insert(lists[x].end(), u.jt.begin(), u.jt.end());
            }
            size_t len = 0;
            for (auto& l : lists) len = std::max(len, l.size() / 2);
            bjob[p][e].assign(2 * len * kXcds, -1);
            for (int x = 0; x < kXcds; x++)
                for (size_t s = 0; s < lists[x].size() / 2; s++) {
                    bjob[p][e][2 * (s * kXcds + x)] = lists[x][2 * s];
                    bjob[p][e][2 * (s * kXcds + x) + 1] = lists[x][2 * s + 1];
                }
            b.tiles2[p][e] = (int)(len * kXcds);
            b.njobs2[p][e] = (int)jobs[p][e].size();
            bytes += align_up(sizeof(GemmJob) * std::max<size_t>(1, jobs[p][e].size()));
            bytes += align_up(sizeof(int32_t) * std::max<size_t>(2, bjob[p][e].size()));
        }
    // FFT engine: (job, first row) of every block of a pass, dealt over
    // every XCD in turn (a pass's work follows the bin count of each design,
    // which the per-design XCD placement of the grids does not balance)
    std::vector<int32_t> fblk[4];
    for (int p = 0; p < 4; p++) {
        for (size_t q = 0; q < fjobs[p].size(); q++)
            for (int r = 0; r < fjobs[p][q].rows; r += fjobs[p][q].rpb) {
                fblk[p].push_back((int32_t)q);
                fblk[p].push_back(r);
            }
        b.nfft[p] = (int)fjobs[p].size();
        b.fft_blocks[p] = (int)(fblk[p].size() / 2);
        bytes += align_up(sizeof(FftJob) * std::max<size_t>(1, fjobs[p].size()));
        bytes += align_up(sizeof(int32_t) * std::max<size_t>(2, fblk[p].size()));
    }
    bytes += align_up(sizeof(PlaneSum) * std::max<size_t>(1, sums.size()));
    // block tables of the six grids
    const std::vector<int32_t>* pfxs[6] = {&pc, &pb, &pw, &ph, &ps, &pt};
    int* grid[6] = {&b.cell_blocks, &b.bin_blocks, &b.wa_blocks, &b.hf_blocks, &b.scat_blocks, &b.tile_blocks};
    if (nd > (1 << kRefBits)) return fail(c, EP_ERR_INVALID_ARG, "too many designs for the block tables");
    const std::vector<int> xcd_of = assign_xcds(ds);
    std::vector<int32_t> blk[6];
    for (int g = 0; g < 6; g++) {
        const std::vector<int32_t>& P = *pfxs[g];
        for (int i = 0; i < nd; i++)
            if (P[i + 1] - P[i] > (1 << (31 - kRefBits)) - 1)
                return fail(c, EP_ERR_INVALID_ARG, "design too large for the block tables");
        build_block_table(P, nd, xcd_of, blk[g]);
        *grid[g] = P[nd] > 0 ? (int)blk[g].size() : 0;
        bytes += align_up(sizeof(int32_t) * blk[g].size());
    }
    if (ext_mem) {
        if (bytes > ext_cap)
            return fail(c, EP_ERR_INVALID_ARG,
                        "launch descriptors of this batch need " + std::to_string(bytes) +
                            " bytes; the preallocated device buffer holds " + std::to_string(ext_cap));
        b.mem = ext_mem;
        b.ext_mem = true;
    } else {
        HIPCHK(c, hipMalloc(&b.mem, bytes));
        b.ext_mem = false;
    }
    std::vector<char> img(bytes, 0);
    size_t off = 0;
    auto put = [&](const void* src, size_t n) {
        size_t o = off;
        memcpy(img.data() + o, src, n);
        off += align_up(n);
        return (char*)b.mem + o;
    };
    b.d_dev = (DesignDev*)put(dv.data(), sizeof(DesignDev) * nd);
    b.bd.d = b.d_dev;
    b.bd.nd = nd;
    b.bd.pfx_cell = (const int32_t*)put(pc.data(), 4 * (nd + 1));
    b.bd.pfx_bin = (const int32_t*)put(pb.data(), 4 * (nd + 1));
    b.bd.pfx_wa = (const int32_t*)put(pw.data(), 4 * (nd + 1));
    b.bd.pfx_hf = (const int32_t*)put(ph.data(), 4 * (nd + 1));
    b.bd.pfx_scat = (const int32_t*)put(ps.data(), 4 * (nd + 1));
    b.bd.pfx_tile = (const int32_t*)put(pt.data(), 4 * (nd + 1));
    const int32_t** blkp[6] = {&b.bd.blk_cell, &b.bd.blk_bin, &b.bd.blk_wa, &b.bd.blk_hf, &b.bd.blk_scat,
                               &b.bd.blk_tile};
    for (int g = 0; g < 6; g++) *blkp[g] = (const int32_t*)put(blk[g].data(), sizeof(int32_t) * blk[g].size());
    const int32_t zero = 0;
    b.dzero = (const int32_t*)put(&zero, 4);
    for (int p = 0; p < 4; p++)
        for (int e = 0; e < kEngines; e++) {
            GemmJob dummy;
            memset(&dummy, 0, sizeof(dummy));
            if (jobs[p][e].empty()) b.jobs2[p][e] = (GemmJob*)put(&dummy, sizeof(GemmJob));
            else b.jobs2[p][e] = (GemmJob*)put(jobs[p][e].data(), sizeof(GemmJob) * jobs[p][e].size());
            const int32_t none[2] = {-1, -1};
            if (bjob[p][e].empty()) b.bjob2[p][e] = (int32_t*)put(none, sizeof(none));
            else b.bjob2[p][e] = (int32_t*)put(bjob[p][e].data(), sizeof(int32_t) * bjob[p][e].size());
        }
    for (int p = 0; p < 4; p++) {
        FftJob dummy;
        memset(&dummy, 0, sizeof(dummy));
        b.fjobs[p] = (FftJob*)(fjobs[p].empty() ? put(&dummy, sizeof(dummy))
                                                : put(fjobs[p].data(), sizeof(FftJob) * fjobs[p].size()));
        const int32_t none[2] = {-1, -1};
        b.fblk[p] = (int32_t*)(fblk[p].empty() ? put(none, sizeof(none))
                                               : put(fblk[p].data(), sizeof(int32_t) * fblk[p].size()));
    }
    {
        PlaneSum none;
        memset(&none, 0, sizeof(none));
        b.sums = (PlaneSum*)(sums.empty() ? put(&none, sizeof(none))
                                          : put(sums.data(), sizeof(PlaneSum) * sums.size()));
        b.n_sums = (int)sums.size();
        b.sum_blocks = 0;
        for (auto& ps : sums) b.sum_blocks = std::max(b.sum_blocks, ps.blocks);
    }
    if (b.ext_mem) {
        // stream-ordered behind any earlier use of the buffer on this stream
        HIPCHK(c, hipMemcpyAsync(b.mem, img.data(), bytes, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));  // img is read until here
    } else {
        HIPCHK(c, hipMemcpy(b.mem, img.data(), bytes, hipMemcpyHostToDevice));
    }
    b.cfgs.assign(nd, ep_config{});
    for (int i = 0; i < nd; i++) b.cfgs[i] = *cfgs[i];
    return EP_OK;
}

int batch_build(Ctx* c, const std::vector<Design*>& ds, const ep_config* cfg, Batch& b) {
    std::vector<const ep_config*> cf(ds.size(), cfg);
    return batch_build_cfgs(c, ds, cf, b, nullptr, 0);
}

void batch_release(Batch& b) {
    for (int k = 0; k < 2; k++) {
        if (b.graph[k]) hipGraphExecDestroy(b.graph[k]);
        b.graph[k] = nullptr;
        b.graph_iters[k] = 0;
    }
    if (b.mem && !b.ext_mem) hipFree(b.mem);
    b.mem = nullptr;
    b.ext_mem = false;
    b.ds.clear();
    for (int p = 0; p < 4; p++) b.nfft[p] = b.fft_blocks[p] = 0;
}

static DevState fresh_state(const Design& D, const ep_config* cfg) {
    DevState s;
    memset(&s, 0, sizeof(s));
    s.lam = 0.0;  // mod.rs:100-101
    s.s[0] = s.s[1] = cfg->nesterov_step_size;
    s.a = 1.0;  // nesterov.rs:30
    s.done = (D.dev.n_mov == 0 || cfg->max_iters == 0) ? 1 : 0;  // mod.rs:80-82
    return s;
}

// Reset the listed designs of a batch (all when idx is empty).
int batch_reset_some(Ctx* c, Batch& b, const std::vector<int>& idx) {
    std::vector<int> which = idx;
    if (which.empty())
        for (int i = 0; i < (int)b.ds.size(); i++) which.push_back(i);
    for (int i : which) {
        Design& D = *b.ds[i];
        c->state_host[D.state_idx] = fresh_state(D, &b.cfgs[i]);
        HIPCHK(c, hipMemcpyAsync(D.dev.st, &c->state_host[D.state_idx], sizeof(DevState),
                                 hipMemcpyHostToDevice, c->stream));
    }
    // positions: a reset launch over the whole batch would disturb running
    // designs, so reset only the listed ones, one small batch each
    if (which.size() == b.ds.size()) {
        launch_reset(b.bd, b.prec, b.cell_blocks, c->stream);
    } else {
        for (int i : which) {
            BatchDev one = b.bd;
            one.d = b.d_dev + i;
            one.nd = 1;
            one.pfx_cell = b.dzero;
            one.blk_cell = nullptr;
            launch_reset(one, b.prec, b.ds[i]->dev.n_cell_blocks, c->stream);
        }
    }
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return EP_OK;
}

int batch_reset(Ctx* c, Batch& b, const ep_config* cfg) {
    (void)cfg;
    return batch_reset_some(c, b, {});
}

static void enqueue_gemm(Ctx* c, Batch& b, int p) {
    for (int e = 0; e < kEngines; e++)
        if (b.njobs2[p][e] > 0)
            launch_gemm_pass(b.jobs2[p][e], b.bjob2[p][e], b.tiles2[p][e], b.d_dev, b.prec, e, p,
                             c->stream);
    if (b.fft_blocks[p] > 0) {
        if (p == 1) launch_fft_solve(b.fjobs[p], b.fblk[p], b.fft_blocks[p], b.d_dev, c->stream);
        else if (p == 3) launch_fft_pair(b.fjobs[p], b.fblk[p], b.fft_blocks[p], c->stream);
        else launch_fft_pass(b.fjobs[p], b.fblk[p], b.fft_blocks[p], b.d_dev, c->stream);
    }
    if (p == 3) launch_plane_sum(b.sums, b.n_sums, b.sum_blocks, c->stream);  // field = sum of planes
}

// Phases of one iteration, in launch order (names used by the profiler).
static const char* kPhaseNames[] = {"sort", "wa", "scatter", "dens_final", "dct_f1", "dct_f2",
                                    "dct_i1", "dct_i2", "gather", "ctrl_a", "fix", "ctrl_b",
                                    "update", "hpwl"};
constexpr int kNumPhases = 14;
constexpr int kWaPhase = 1;
constexpr int kGatherPhase = 8;
constexpr int kCtrlPhase = 9;
constexpr int kFixPhase = 10;
constexpr int kCtrlEndPhase = 11;

static void enqueue_phase(Ctx* c, Batch& b, int ph) {
    hipStream_t s = c->stream;
    switch (ph) {
        case 0: launch_tile_sort(b.bd, b.prec, b.cell_blocks, s); break;      // scatter order
        case 1:  // mod.rs:110-115
            launch_wa(b.bd, b.prec, b.wa_blocks, b.hf_blocks, s);
            launch_wl_cell(b.bd, b.prec, b.cell_blocks, s);
            break;
        case 2:                                                               // mod.rs:118-119
            launch_scatter_pull(b.bd, b.prec, b.tile_blocks, s);
            launch_scatter(b.bd, b.prec, b.scat_blocks, s);
            break;
        case 3: launch_dens_final(b.bd, b.prec, b.bin_blocks, s); break;      // mod.rs:119
        case 4: case 5: case 6: case 7: enqueue_gemm(c, b, ph - 4); break;   // mod.rs:121-123
        case 8: launch_gather(b.bd, b.prec, b.cell_blocks, 0, s); break;      // mod.rs:126-156
        case 9: launch_ctrl(b.bd, 0, s); break;                               // mod.rs:131-143, 159-166
        case 10: launch_gather(b.bd, b.prec, b.cell_blocks, 1, s); break;
        case 11: launch_ctrl(b.bd, 1, s); break;                              // nesterov.rs:41-47, mod.rs:176-195
        case 12: launch_update(b.bd, b.prec, b.cell_blocks, s); break;        // mod.rs:169-173, 105-107
        case 13: if (b.log_hpwl) launch_hpwl(b.bd, b.prec, b.wa_blocks, b.hf_blocks, 0, s); break;
    }
}

// One iteration of mod.rs:104-197 for every live design of the batch.  The
// WA gradient depends only on the look-ahead positions, as the density chain
// (sort, scatter, transforms) does, and the two meet at the gather: the WA
// kernels run on the side stream, forked after the previous update and
// joined before the gather, so they fill the CUs the latency-bound density
// chain leaves idle (in a captured chunk the fork and join become graph
// edges).  The fix pass (combine once the penalty was initialised by ctrl_a
// in the same iteration) can only do work while some design's penalty is
// unset.
static void enqueue_iteration(Ctx* c, Batch& b, bool with_fix) {
    hipEventRecord(c->ev_fork, c->stream);
    // the density chain's first kernels are enqueued (captured) before the
    // side stream's, so the chain does not wait behind the fork
    enqueue_phase(c, b, 0);
    hipStreamWaitEvent(c->side, c->ev_fork, 0);
    launch_wa(b.bd, b.prec, b.wa_blocks, b.hf_blocks, c->side);
    launch_wl_cell(b.bd, b.prec, b.cell_blocks, c->side);
    hipEventRecord(c->ev_join, c->side);
    for (int ph = 1; ph < kNumPhases; ph++) {
        if (ph == kWaPhase) continue;
        if (!with_fix && (ph == kFixPhase || ph == kCtrlEndPhase)) continue;
        if (ph == kGatherPhase) hipStreamWaitEvent(c->stream, c->ev_join, 0);
        if (ph == kCtrlPhase && !with_fix)
            launch_ctrl(b.bd, 4, c->stream);  // both control steps, nothing between them
        else
            enqueue_phase(c, b, ph);
    }
}

// True while a live design of the batch has no penalty yet, from the host
// mirror of the states (fresh after every reset and every read_states), so
// a chunk launched without the fix pass never needs it: the penalty is set
// once and never unset.
static bool batch_needs_fix(const Ctx* c, const Batch& b) {
    for (const Design* D : b.ds) {
        const DevState& s = c->state_host[D->state_idx];
        if (!s.done && !s.lam_init) return true;
    }
    return false;
}

// Makes the batch's graph of `iters` iterations (with or without the fix
// pass) current: captured and instantiated, or, when a kept graph of the
// same launches exists, captured and updated in place.  Launches nothing, so
// a caller can prepare the next chunk's graph while the GPU runs this one.
int batch_graph_ready(Ctx* c, Batch& b, int iters, bool with_fix) {
    const int k = with_fix ? 1 : 0;
    hipGraphExec_t& ge = b.graph[k];
    int& gi = b.graph_iters[k];
    if (ge && gi == iters && !b.graph_stale[k]) return EP_OK;
    hipGraph_t g = nullptr;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < iters; i++) enqueue_iteration(c, b, with_fix);
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (e != hipSuccess) return fail(c, EP_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(e));
    bool updated = false;
    if (ge && gi == iters && b.graph_stale[k]) {
        // same launches, new grids and arguments: update the instantiated graph
        hipGraphNode_t bad = nullptr;
        hipGraphExecUpdateResult res = hipGraphExecUpdateError;
        updated = hipGraphExecUpdate(ge, g, &bad, &res) == hipSuccess && res == hipGraphExecUpdateSuccess;
        if (!updated) (void)hipGetLastError();  // the launches changed: instantiate afresh
    }
    if (updated) {
        c->graph_updates++;
    } else {
        if (ge) hipGraphExecDestroy(ge);
        ge = nullptr;
        e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            hipGraphDestroy(g);
            return fail(c, EP_ERR_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(e));
        }
        c->graph_captures++;
    }
    hipGraphDestroy(g);
    gi = iters;
    b.graph_stale[k] = false;
    return EP_OK;
}

int batch_run_iters(Ctx* c, Batch& b, int iters, bool use_graph, int fix) {
    if (b.ds.empty() || iters <= 0) return EP_OK;
    const bool with_fix = fix < 0 ? batch_needs_fix(c, b) : fix != 0;
    if (!use_graph) {
        for (int i = 0; i < iters; i++) enqueue_iteration(c, b, with_fix);
        HIPCHK(c, hipGetLastError());
        return EP_OK;
    }
    int rc = batch_graph_ready(c, b, iters, with_fix);
    if (rc) return rc;
    HIPCHK(c, hipGraphLaunch(b.graph[with_fix ? 1 : 0], c->stream));
    return EP_OK;
}

int read_states(Ctx* c, const Batch& b) {
    if (b.ds.empty()) return EP_OK;
    int lo = kStatePool, hi = -1;
    for (auto* D : b.ds) {
        lo = std::min(lo, D->state_idx);
        hi = std::max(hi, D->state_idx);
    }
    HIPCHK(c, hipMemcpyAsync(&c->state_host[lo], &c->state_pool[lo], sizeof(DevState) * (hi - lo + 1),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return EP_OK;
}

int batch_finalize(Ctx* c, Batch& b) {
    launch_ctrl(b.bd, 2, c->stream);
    launch_hpwl(b.bd, b.prec, b.wa_blocks, b.hf_blocks, 1, c->stream);
    launch_ctrl(b.bd, 3, c->stream);
    HIPCHK(c, hipGetLastError());
    return read_states(c, b);
}

void fill_result(const Design& D, const DevState& s, ep_result* r) {
    (void)D;
    r->status = s.nonfinite ? EP_ERR_NONFINITE : EP_OK;
    r->iters = s.it;
    r->converged = s.converged;
    r->precision = D.prec;
    r->overflow = s.ov;
    r->hpwl = s.final_hpwl;
    r->lambda = s.lam;
}

// Runs a built batch until every design stops.  The next chunk of
// iterations is queued before the host waits for the states copied behind
// the current one, so the GPU does not idle while the host checks them,
// except where the states say the batch is about to finish.  Each design's time is taken on the device, from the reset to the
// control step that stopped it, so it does not depend on when the host looks.
int batch_place(Ctx* c, Batch& b, int sync_every, bool use_graph) {
    int rc = batch_reset(c, b, nullptr);
    if (rc) return rc;
    int max_it = 0;
    for (auto& cf : b.cfgs) max_it = std::max(max_it, cf.max_iters);
    b.done_ms.assign(b.ds.size(), -1.0);
    int lo = kStatePool, hi = -1;
    for (auto* D : b.ds) {
        lo = std::min(lo, D->state_idx);
        hi = std::max(hi, D->state_idx);
    }
    auto all_done = [&]() {
        for (const Design* D : b.ds)
            if (!c->state_host[D->state_idx].done) return false;
        return true;
    };
    auto copy_states = [&]() -> int {
        HIPCHK(c, hipMemcpyAsync(&c->state_host[lo], &c->state_pool[lo], sizeof(DevState) * (hi - lo + 1),
                                 hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipEventRecord(c->ev_states, c->stream));
        return EP_OK;
    };
    int queued = 0;
    // Once every design has stopped, the chunk queued behind the one whose
    // states the host reads runs iterations that skip every design, and each
    // still costs a launch per kernel (about 35 us on the 64-design batch).
    // When every live design is past its minimum iteration count within the
    // next chunk and within 5 % of its overflow target, that chunk will most
    // likely stop them all, so nothing is queued behind it; a wrong guess
    // costs one host round trip with the GPU idle.
    auto ending = [&]() {
        for (size_t k = 0; k < b.ds.size(); k++) {
            const DevState& s = c->state_host[b.ds[k]->state_idx];
            if (s.done) continue;
            const ep_config& cf = b.cfgs[k];
            if (s.it + sync_every - 1 < cf.min_iters || !(s.ov < cf.overflow_target * 1.05)) return false;
        }
        return true;
    };
    // a penalty once set stays set, so a mirror one chunk old can only ask
    // for the fix pass where it is no longer needed (it then does nothing)
    bool fix = batch_needs_fix(c, b);
    bool ahead = true;
    if (!all_done() && max_it > 0) {
        if ((rc = batch_run_iters(c, b, sync_every, use_graph, fix))) return rc;
        queued += sync_every;
        if ((rc = copy_states())) return rc;
        for (;;) {
            const bool queued_ahead = ahead && queued < max_it;
            if (queued_ahead) {
                if ((rc = batch_run_iters(c, b, sync_every, use_graph, fix))) return rc;
                queued += sync_every;
            }
            HIPCHK(c, hipEventSynchronize(c->ev_states));
            if (all_done() || queued >= max_it) break;
            fix = batch_needs_fix(c, b);
            ahead = !ending();
            if (!queued_ahead) {  // the guess was wrong: the next chunk now
                if ((rc = batch_run_iters(c, b, sync_every, use_graph, fix))) return rc;
                queued += sync_every;
            }
            if ((rc = copy_states())) return rc;
        }
    }
    rc = batch_finalize(c, b);
    for (size_t k = 0; k < b.ds.size(); k++) {
        const DevState& s = c->state_host[b.ds[k]->state_idx];
        b.done_ms[k] = (s.t_done > s.t_start && c->wall_khz > 0.0)
                           ? (double)(s.t_done - s.t_start) / c->wall_khz : 0.0;
    }
    return rc;
}

Design* find_design(Ctx* c, uint32_t id) {
    auto it = c->designs.find(id);
    return it == c->designs.end() ? nullptr : it->second.get();
}

// ------------------------------------------------------------ caching ----

Batch* cached_batch(Ctx* c, const std::vector<Design*>& ds, const ep_config* cfg, int* rc) {
    *rc = EP_OK;
    if (c->cache && c->cache->ds == ds && c->cache_cfg_valid &&
        memcmp(&c->cache_cfg, cfg, sizeof(ep_config)) == 0)
        return c->cache.get();
    release_batch_cache(c);
    c->cache.reset(new Batch());
    *rc = batch_build(c, ds, cfg, *c->cache);
    if (*rc) {
        release_batch_cache(c);
        return nullptr;
    }
    c->cache_cfg = *cfg;
    c->cache_cfg_valid = true;
    return c->cache.get();
}

void release_batch_cache(Ctx* c) {
    if (c->cache) batch_release(*c->cache);
    c->cache.reset();
    c->cache_cfg_valid = false;
}

int free_design_locked(Ctx* c, uint32_t id) {
    auto it = c->designs.find(id);
    if (it == c->designs.end()) return fail(c, EP_ERR_NOT_FOUND, "unknown design id");
    if (c->cache)
        for (auto* D : c->cache->ds)
            if (D->id == id) { release_batch_cache(c); break; }
    hipStreamSynchronize(c->stream);
    arena_release(c, it->second->arena, it->second->bytes);
    c->used -= it->second->bytes;
    c->free_states.push_back(it->second->state_idx);
    c->designs.erase(it);
    return EP_OK;
}

}  // namespace ep

using namespace ep;

static bool use_graphs() { return getenv("EP_NO_GRAPH") == nullptr; }

extern "C" {

int ep_design_free(ep_ctx* h, uint32_t id) {
    Ctx* c = (Ctx*)h;
    if (!c) return fail(nullptr, EP_ERR_INVALID_ARG, "null context");
    std::lock_guard<std::recursive_mutex> g(c->mu);
    return free_design_locked(c, id);
}

int ep_place_batch(ep_ctx* h, const uint32_t* ids, uint32_t n, const ep_config* cfg, ep_result* res) {
    Ctx* c = (Ctx*)h;
    if (!c || (!ids && n) || !cfg || (!res && n)) return fail(c, EP_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::recursive_mutex> g(c->mu);
    if (c->server_running) return fail(c, EP_ERR_INVALID_ARG, "context is serving; stop it first");
    HIPCHK(c, hipSetDevice(c->device));
    std::vector<Design*> ds;
    for (uint32_t i = 0; i < n; i++) {
        Design* D = find_design(c, ids[i]);
        if (!D) return fail(c, EP_ERR_NOT_FOUND, "unknown design id");
        int rc = validate_config(c, cfg, D->log_cap);
        if (rc) return rc;
        ds.push_back(D);
    }
    if (ds.empty()) return EP_OK;
    int rc;
    Batch* b = cached_batch(c, ds, cfg, &rc);
    if (!b) return rc;
    const int sync_every = cfg->legalize_interval > 0 ? cfg->legalize_interval : 5;
    rc = batch_place(c, *b, sync_every, use_graphs());
    if (rc) return rc;
    int worst = EP_OK;
    for (uint32_t i = 0; i < n; i++) {
        fill_result(*ds[i], c->state_host[ds[i]->state_idx], &res[i]);
        res[i].t_submit_ms = 0.0;
        res[i].t_admit_ms = 0.0;
        res[i].t_done_ms = i < b->done_ms.size() ? b->done_ms[i] : 0.0;
        if (res[i].status != EP_OK) worst = res[i].status;
    }
    if (worst == EP_ERR_NONFINITE) fail(c, worst, "non-finite value in the placement state");
    return worst;
}

int ep_place(ep_ctx* h, uint32_t id, const ep_config* cfg, ep_result* res) {
    return ep_place_batch(h, &id, 1, cfg, res);
}

int ep_positions(ep_ctx* h, uint32_t id, double* x, double* y) {
    Ctx* c = (Ctx*)h;
    if (!c || !x || !y) return fail(c, EP_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::recursive_mutex> g(c->mu);
    Design* D = find_design(c, id);
    if (!D) return fail(c, EP_ERR_NOT_FOUND, "unknown design id");
    const size_t N = (size_t)D->dev.n_mov;
    if (N == 0) return EP_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int ax = 0; ax < 2; ax++) {
        double* out = ax ? y : x;
        if (D->prec == EP_PREC_FP64) {
            HIPCHK(c, hipMemcpy(out, D->dev.pos[ax], N * 8, hipMemcpyDeviceToHost));
        } else {
            std::vector<float> f(N);
            HIPCHK(c, hipMemcpy(f.data(), D->dev.pos[ax], N * 4, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < N; i++) out[i] = f[i];
        }
    }
    return EP_OK;
}

int ep_iter_log(ep_ctx* h, uint32_t id, ep_iter_rec* recs, uint32_t cap, uint32_t* n) {
    Ctx* c = (Ctx*)h;
    if (!c || !n || (!recs && cap)) return fail(c, EP_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::recursive_mutex> g(c->mu);
    Design* D = find_design(c, id);
    if (!D) return fail(c, EP_ERR_NOT_FOUND, "unknown design id");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    DevState s;
    HIPCHK(c, hipMemcpy(&s, D->dev.st, sizeof(DevState), hipMemcpyDeviceToHost));
    uint32_t k = (uint32_t)std::max(0, std::min(s.n_log, D->log_cap));
    *n = k;
    k = std::min(k, cap);
    if (k) HIPCHK(c, hipMemcpy(recs, D->dev.log, k * sizeof(ep_iter_rec), hipMemcpyDeviceToHost));
    return EP_OK;
}

}  // extern "C"

// ---------------------------------------------------------- autotuner ----

// Chooses the transform engine of every distinct bin shape among `ids` by
// timing the transform passes of the whole batch `ids` (the launches the
// placement loop makes, so each candidate is measured under the occupancy
// the other shapes' tiles impose): first every new shape on one engine (MFMA,
// FFT, MFMA K64, where eligible), the fastest kept as the start; then one
// shape at a time from the largest, each candidate with the other shapes at
// their current choice; the fastest is kept in the context's table.  With a cache path, shapes
// already in the file are taken from it and new results are appended.
// Engines differ only in FP32 rounding order, and the table is per shape,
// so a design's results do not depend on the batch it runs in.
static int time_transforms(Ctx* c, const std::vector<Design*>& ds, uint32_t reps, float* ms_out) {
    ep_config cfg;
    ep_config_default(&cfg);
    cfg.field_mode = EP_FIELD_IDXST;
    release_batch_cache(c);
    Batch b;
    std::vector<const ep_config*> cf(ds.size(), &cfg);
    int rc = batch_build_cfgs(c, ds, cf, b, nullptr, 0);
    if (rc) return rc;
    // live design states so the transform kernels do not skip them
    for (Design* D : ds) {
        DevState st;
        memset(&st, 0, sizeof(st));
        st.a = 1.0;
        c->state_host[D->state_idx] = st;
        hipError_t er = hipMemcpyAsync(D->dev.st, &c->state_host[D->state_idx], sizeof(DevState),
                                       hipMemcpyHostToDevice, c->stream);
        if (er != hipSuccess) { batch_release(b); return fail(c, EP_ERR_DEVICE, hipGetErrorString(er)); }
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t er = hipEventCreate(&e0);
    if (er == hipSuccess) er = hipEventCreate(&e1);
    if (er == hipSuccess) {
        for (int p = 0; p < 4; p++) enqueue_gemm(c, b, p);  // warm
        hipEventRecord(e0, c->stream);
        for (uint32_t r = 0; r < reps; r++)
            for (int p = 0; p < 4; p++) enqueue_gemm(c, b, p);
        hipEventRecord(e1, c->stream);
        er = hipEventSynchronize(e1);
    }
    float ms = 0.f;
    if (er == hipSuccess) er = hipEventElapsedTime(&ms, e0, e1);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    batch_release(b);
    if (er != hipSuccess) return fail(c, EP_ERR_DEVICE, hipGetErrorString(er));
    *ms_out = ms;
    return EP_OK;
}

extern "C" int ep_autotune(ep_ctx* h, const uint32_t* ids, uint32_t n, uint32_t reps,
                           const char* cache_path) {
    Ctx* c = (Ctx*)h;
    if (!c || (!ids && n)) return fail(c, EP_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::recursive_mutex> g(c->mu);
    if (c->server_running) return fail(c, EP_ERR_INVALID_ARG, "context is serving; stop it first");
    HIPCHK(c, hipSetDevice(c->device));
    std::map<std::pair<int, int>, int> cached;
    if (cache_path) {
        FILE* f = fopen(cache_path, "r");
        if (f) {
            int bx, by, e;
            while (fscanf(f, "%d %d %d", &bx, &by, &e) == 3)
                if (e >= 0 && e < kNumEngines) {
                    const std::vector<int> ok = engines_for_shape(bx, by);
                    if (std::find(ok.begin(), ok.end(), e) != ok.end()) cached[{bx, by}] = e;
                }
            fclose(f);
        }
    }
    if (reps == 0) reps = 5;
    std::vector<Design*> ds;
    std::set<std::pair<int, int>> shapes;
    for (uint32_t i = 0; i < n; i++) {
        Design* D = find_design(c, ids[i]);
        if (!D) return fail(c, EP_ERR_NOT_FOUND, "unknown design id");
        if (D->prec != EP_PREC_FAST) continue;
        if (std::find(ds.begin(), ds.end(), D) == ds.end()) ds.push_back(D);
        if (engines_for_shape(D->dev.Bx, D->dev.By).size() > 1) shapes.insert({D->dev.Bx, D->dev.By});
    }
    std::vector<std::pair<int, int>> todo;
    for (auto& sh : shapes) {
        auto hit = cached.find(sh);
        if (hit != cached.end()) c->tune[sh] = hit->second;
        else todo.push_back(sh);
    }
    // largest shapes first: they bound a batched pass
    std::sort(todo.begin(), todo.end(), [](const std::pair<int, int>& a, const std::pair<int, int>& b) {
        return (int64_t)a.first * a.second > (int64_t)b.first * b.second;
    });
    for (auto& sh : todo)
        if (!c->tune.count(sh)) c->tune[sh] = engine_for(c, EP_PREC_FAST, sh.first, sh.second);
    // uniform starting points first: every new shape on one engine where it
    // is eligible (a pass whose shapes take two engines makes two launches,
    // which a shape-by-shape search starting from one engine never leaves)
    if (!todo.empty()) {
        std::map<std::pair<int, int>, int> start = c->tune;
        float best = 1e30f;
        for (int e : {0, kEngineFft, 3}) {
            std::map<std::pair<int, int>, int> trial = start;
            bool any = false;
            for (auto& sh : todo) {
                const std::vector<int> ok = engines_for_shape(sh.first, sh.second);
                if (std::find(ok.begin(), ok.end(), e) != ok.end()) { trial[sh] = e; any = true; }
            }
            if (!any) continue;
            c->tune = trial;
            float ms = 0.f;
            int rc = time_transforms(c, ds, reps, &ms);
            if (rc) return rc;
            if (ms < best) { best = ms; start = trial; }
        }
        c->tune = start;
    }
    std::vector<std::pair<std::pair<int, int>, int>> fresh;
    for (auto& sh : todo) {
        float best = 1e30f;
        int best_e = c->tune[sh];
        const bool splittable = sh.first >= kSplitMinK || sh.second >= kSplitMinK;
        for (int e : engines_for_shape(sh.first, sh.second)) {
            if (e == 4 && !splittable) continue;  // same launches as engine 0
            c->tune[sh] = e;
            float ms = 0.f;
            int rc = time_transforms(c, ds, reps, &ms);
            if (rc) return rc;
            if (ms < best) { best = ms; best_e = e; }
        }
        c->tune[sh] = best_e;
        fresh.push_back({sh, best_e});
    }
    release_batch_cache(c);
    if (cache_path && !fresh.empty()) {
        FILE* f = fopen(cache_path, "a");
        if (f) {
            for (auto& r : fresh) fprintf(f, "%d %d %d\n", r.first.first, r.first.second, r.second);
            fclose(f);
        }
    }
    return EP_OK;
}

// Reads the tuning table: up to cap (Bx, By, engine) triples.
extern "C" int ep_tune_table(ep_ctx* h, int32_t* out, uint32_t cap, uint32_t* n) {
    Ctx* c = (Ctx*)h;
    if (!c || !n || (!out && cap)) return fail(c, EP_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::recursive_mutex> g(c->mu);
    uint32_t k = 0;
    for (auto& kv : c->tune) {
        if (k < cap) {
            out[3 * k] = kv.first.first;
            out[3 * k + 1] = kv.first.second;
            out[3 * k + 2] = kv.second;
        }
        k++;
    }
    *n = k;
    return EP_OK;
}

// Pins the transform engine of one bin shape (engine < 0 removes the entry).
extern "C" int ep_tune_set(ep_ctx* h, int32_t bx, int32_t by, int32_t engine) {
    Ctx* c = (Ctx*)h;
    if (!c) return fail(c, EP_ERR_INVALID_ARG, "null context");
    if (engine >= kNumEngines) return fail(c, EP_ERR_INVALID_ARG, "unknown transform engine");
    if (engine == kEngineFft && !(fft_len_ok(bx) && fft_len_ok(by)))
        return fail(c, EP_ERR_INVALID_ARG, "the FFT engine needs bin counts that are powers of two from 16 to 512");
    if (engine >= 0 && engine != 1 && engine != kEngineFft && (bx % 64 || by % 64))
        return fail(c, EP_ERR_INVALID_ARG, "MFMA engines need bins that are multiples of 64");
    std::lock_guard<std::recursive_mutex> g(c->mu);
    if (c->server_running) return fail(c, EP_ERR_INVALID_ARG, "context is serving; stop it first");
    if (engine < 0) c->tune.erase({bx, by});
    else c->tune[{bx, by}] = engine;
    release_batch_cache(c);  // cached batches hold transform jobs split by engine
    return EP_OK;
}

// ---------------------------------------------------------- profiling ----

extern "C" int ep_calibrate_copy(ep_ctx* h, uint64_t bytes) {
    Ctx* c = (Ctx*)h;
    if (!c || bytes == 0) return fail(c, EP_ERR_INVALID_ARG, "bad argument");
    std::lock_guard<std::recursive_mutex> g(c->mu);
    HIPCHK(c, hipSetDevice(c->device));
    void *a = nullptr, *b = nullptr;
    HIPCHK(c, hipMalloc(&a, bytes));
    hipError_t e = hipMalloc(&b, bytes);
    if (e == hipSuccess) e = hipMemsetAsync(a, 1, bytes, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    hipFree(a);
    if (b) hipFree(b);
    if (e != hipSuccess) return fail(c, EP_ERR_DEVICE, hipGetErrorString(e));
    return EP_OK;
}

extern "C" int ep_profile_iteration(ep_ctx* h, const uint32_t* ids, uint32_t n, const ep_config* cfg,
                                    uint32_t warm, uint32_t reps, ep_phase_time* out, uint32_t cap,
                                    uint32_t* n_out) {
    Ctx* c = (Ctx*)h;
    if (!c || !ids || !cfg || !n_out || (!out && cap)) return fail(c, EP_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::recursive_mutex> g(c->mu);
    if (c->server_running) return fail(c, EP_ERR_INVALID_ARG, "context is serving; stop it first");
    HIPCHK(c, hipSetDevice(c->device));
    std::vector<Design*> ds;
    for (uint32_t i = 0; i < n; i++) {
        Design* D = find_design(c, ids[i]);
        if (!D) return fail(c, EP_ERR_NOT_FOUND, "unknown design id");
        ds.push_back(D);
    }
    int rc;
    Batch* b = cached_batch(c, ds, cfg, &rc);
    if (!b) return rc;
    if ((rc = batch_reset(c, *b, cfg))) return rc;
    if ((rc = batch_run_iters(c, *b, (int)warm, false))) return rc;
    std::vector<hipEvent_t> ev(kNumPhases + 1);
    for (auto& e : ev) HIPCHK(c, hipEventCreate(&e));
    std::vector<double> acc(kNumPhases, 0.0);
    for (uint32_t r = 0; r < reps; r++) {
        HIPCHK(c, hipEventRecord(ev[0], c->stream));
        for (int ph = 0; ph < kNumPhases; ph++) {
            enqueue_phase(c, *b, ph);
            HIPCHK(c, hipEventRecord(ev[ph + 1], c->stream));
        }
        HIPCHK(c, hipEventSynchronize(ev[kNumPhases]));
        for (int ph = 0; ph < kNumPhases; ph++) {
            float ms = 0.f;
            HIPCHK(c, hipEventElapsedTime(&ms, ev[ph], ev[ph + 1]));
            acc[ph] += ms;
        }
    }
    for (auto& e : ev) hipEventDestroy(e);
    *n_out = kNumPhases;
    for (uint32_t ph = 0; ph < (uint32_t)kNumPhases && ph < cap; ph++) {
        memset(&out[ph], 0, sizeof(ep_phase_time));
        strncpy(out[ph].name, kPhaseNames[ph], sizeof(out[ph].name) - 1);
        out[ph].avg_ms = reps ? acc[ph] / reps : 0.0;
        out[ph].launches = reps;
    }
    return EP_OK;
}

// -------------------------------------------------------- evaluation -----

namespace ep {

// A single-design batch for the evaluation calls (context eval config).
static int eval_prep(Ctx* c, uint32_t id, Design** Dout, Batch** bout) {
    Design* D = find_design(c, id);
    if (!D) return fail(c, EP_ERR_NOT_FOUND, "unknown design id");
    HIPCHK(c, hipSetDevice(c->device));
    std::vector<Design*> ds{D};
    int rc;
    Batch* b = cached_batch(c, ds, &c->eval_cfg, &rc);
    if (!b) return rc;
    DevState s;
    memset(&s, 0, sizeof(s));
    s.lam = 0.0;
    s.a = 1.0;
    c->state_host[D->state_idx] = s;
    HIPCHK(c, hipMemcpyAsync(D->dev.st, &c->state_host[D->state_idx], sizeof(DevState),
                             hipMemcpyHostToDevice, c->stream));
    *Dout = D;
    *bout = b;
    return EP_OK;
}

static int put_T(Ctx* c, void* dst, const double* src, size_t n, int prec) {
    if (!n) return EP_OK;
    if (prec == EP_PREC_FP64) {
        HIPCHK(c, hipMemcpyAsync(dst, src, n * 8, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    } else {
        std::vector<float> f(n);
        for (size_t i = 0; i < n; i++) f[i] = (float)src[i];
        HIPCHK(c, hipMemcpyAsync(dst, f.data(), n * 4, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return EP_OK;
}

// x and y interleaved into a T [n][2] device array.
static int put_xy(Ctx* c, void* dst, const double* x, const double* y, size_t n, int prec) {
    std::vector<double> xy(2 * n);
    for (size_t i = 0; i < n; i++) {
        xy[2 * i] = x[i];
        xy[2 * i + 1] = y[i];
    }
    return put_T(c, dst, xy.data(), 2 * n, prec);
}

static int get_T(Ctx* c, double* dst, const void* src, size_t n, int prec) {
    if (!n) return EP_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (prec == EP_PREC_FP64) {
        HIPCHK(c, hipMemcpy(dst, src, n * 8, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> f(n);
        HIPCHK(c, hipMemcpy(f.data(), src, n * 4, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; i++) dst[i] = f[i];
    }
    return EP_OK;
}

}  // namespace ep

extern "C" {

int ep_eval_config(ep_ctx* h, const ep_config* cfg) {
    Ctx* c = (Ctx*)h;
    if (!c || !cfg) return fail(c, EP_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::recursive_mutex> g(c->mu);
    int rc = validate_config(c, cfg, kDefaultLogCap);
    if (rc) return rc;
    c->eval_cfg = *cfg;
    release_batch_cache(c);
    return EP_OK;
}

int ep_eval_wa_grad(ep_ctx* h, uint32_t id, const double* x, const double* y, double* gx, double* gy) {
    Ctx* c = (Ctx*)h;
    if (!c || !x || !y || !gx || !gy) return fail(c, EP_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::recursive_mutex> g(c->mu);
    Design* D;
    Batch* b;
    int rc = eval_prep(c, id, &D, &b);
    if (rc) return rc;
    const size_t N = (size_t)D->dev.n_mov;
    if ((rc = put_xy(c, D->dev.look, x, y, N, D->prec))) return rc;
    if (!N) return EP_OK;
    double* dbuf = nullptr;
    HIPCHK(c, hipMalloc(&dbuf, N * 16));
    launch_wa(b->bd, b->prec, b->wa_blocks, b->hf_blocks, c->stream);
    launch_wa_reduce(b->bd, b->prec, b->cell_blocks, dbuf, dbuf + N, c->stream);
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(gx, dbuf, N * 8, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(gy, dbuf + N, N * 8, hipMemcpyDeviceToHost);
    hipFree(dbuf);
    if (e != hipSuccess) return fail(c, EP_ERR_DEVICE, hipGetErrorString(e));
    return EP_OK;
}

int ep_eval_density(ep_ctx* h, uint32_t id, const double* x, const double* y, double* dens,
                    double* overflow) {
    Ctx* c = (Ctx*)h;
    if (!c || !x || !y || !dens) return fail(c, EP_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::recursive_mutex> g(c->mu);
    Design* D;
    Batch* b;
    int rc = eval_prep(c, id, &D, &b);
    if (rc) return rc;
    const size_t N = (size_t)D->dev.n_mov, NB = (size_t)D->dev.Bx * D->dev.By;
    if ((rc = put_xy(c, D->dev.look, x, y, N, D->prec))) return rc;
    launch_tile_sort(b->bd, b->prec, b->cell_blocks, c->stream);
    launch_scatter_pull(b->bd, b->prec, b->tile_blocks, c->stream);
    launch_scatter(b->bd, b->prec, b->scat_blocks, c->stream);
    launch_dens_final(b->bd, b->prec, b->bin_blocks, c->stream);
    HIPCHK(c, hipGetLastError());
    if ((rc = get_T(c, dens, D->dev.dens, NB, D->prec))) return rc;
    if (overflow) {
        std::vector<double> p((size_t)D->dev.n_ov_parts * kBinParts);
        HIPCHK(c, hipMemcpy(p.data(), D->dev.part_bin, p.size() * 8, hipMemcpyDeviceToHost));
        double tot = 0, ov = 0;
        for (int k = 0; k < D->dev.n_ov_parts; k++) {
            tot += p[(size_t)k * kBinParts];
            ov += p[(size_t)k * kBinParts + 1];
        }
        *overflow = tot < 1e-30 ? 0.0 : ov / tot;
    }
    return EP_OK;
}

int ep_eval_bin_ranges(ep_ctx* h, uint32_t id, const double* x, const double* y, int32_t* ranges) {
    Ctx* c = (Ctx*)h;
    if (!c || !x || !y || !ranges) return fail(c, EP_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::recursive_mutex> g(c->mu);
    Design* D;
    Batch* b;
    int rc = eval_prep(c, id, &D, &b);
    if (rc) return rc;
    const size_t N = (size_t)D->dev.n_mov;
    if (!N) return EP_OK;
    if ((rc = put_xy(c, D->dev.look, x, y, N, D->prec))) return rc;
    int32_t* dbuf = nullptr;
    HIPCHK(c, hipMalloc(&dbuf, N * 16));
    launch_bin_ranges(b->bd, b->cell_blocks, dbuf, c->stream);
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(ranges, dbuf, N * 16, hipMemcpyDeviceToHost);
    hipFree(dbuf);
    if (e != hipSuccess) return fail(c, EP_ERR_DEVICE, hipGetErrorString(e));
    return EP_OK;
}

int ep_eval_field(ep_ctx* h, uint32_t id, const double* dens, double* ex, double* ey) {
    Ctx* c = (Ctx*)h;
    if (!c || !dens || !ex || !ey) return fail(c, EP_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::recursive_mutex> g(c->mu);
    Design* D;
    Batch* b;
    int rc = eval_prep(c, id, &D, &b);
    if (rc) return rc;
    const size_t NB = (size_t)D->dev.Bx * D->dev.By;
    if ((rc = put_T(c, D->dev.dens, dens, NB, D->prec))) return rc;
    for (int p = 0; p < 4; p++) enqueue_gemm(c, *b, p);
    HIPCHK(c, hipGetLastError());
    std::vector<double> xy(2 * NB);
    if ((rc = get_T(c, xy.data(), D->dev.field, 2 * NB, D->prec))) return rc;
    for (size_t k = 0; k < NB; k++) {
        ex[k] = xy[2 * k];
        ey[k] = xy[2 * k + 1];
    }
    return EP_OK;
}

int ep_eval_density_grad(ep_ctx* h, uint32_t id, const double* x, const double* y, const double* ex,
                         const double* ey, double* gx, double* gy) {
    Ctx* c = (Ctx*)h;
    if (!c || !x || !y || !ex || !ey || !gx || !gy) return fail(c, EP_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::recursive_mutex> g(c->mu);
    Design* D;
    Batch* b;
    int rc = eval_prep(c, id, &D, &b);
    if (rc) return rc;
    const size_t N = (size_t)D->dev.n_mov, NB = (size_t)D->dev.Bx * D->dev.By;
    if (!N) return EP_OK;
    if ((rc = put_xy(c, D->dev.look, x, y, N, D->prec))) return rc;
    if ((rc = put_xy(c, D->dev.field, ex, ey, NB, D->prec))) return rc;
    double* dbuf = nullptr;
    HIPCHK(c, hipMalloc(&dbuf, N * 16));
    launch_gather_export(b->bd, b->prec, b->cell_blocks, dbuf, dbuf + N, c->stream);
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(gx, dbuf, N * 8, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(gy, dbuf + N, N * 8, hipMemcpyDeviceToHost);
    hipFree(dbuf);
    if (e != hipSuccess) return fail(c, EP_ERR_DEVICE, hipGetErrorString(e));
    return EP_OK;
}

int ep_eval_hpwl(ep_ctx* h, uint32_t id, const double* x, const double* y, double* hpwl) {
    Ctx* c = (Ctx*)h;
    if (!c || !x || !y || !hpwl) return fail(c, EP_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::recursive_mutex> g(c->mu);
    Design* D;
    Batch* b;
    int rc = eval_prep(c, id, &D, &b);
    if (rc) return rc;
    const size_t N = (size_t)D->dev.n_mov;
    if ((rc = put_T(c, D->dev.pos[0], x, N, D->prec))) return rc;
    if ((rc = put_T(c, D->dev.pos[1], y, N, D->prec))) return rc;
    launch_hpwl(b->bd, b->prec, b->wa_blocks, b->hf_blocks, 1, c->stream);
    launch_ctrl(b->bd, 3, c->stream);
    HIPCHK(c, hipGetLastError());
    if ((rc = read_states(c, *b))) return rc;
    *hpwl = c->state_host[D->state_idx].final_hpwl;
    return EP_OK;
}

}  // extern "C"
