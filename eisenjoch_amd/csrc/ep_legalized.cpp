// place_electro end to end (crates/nextpnr/src/placer/electro_place/mod.rs:65-204):
// the continuous loop on the device in chunks that end on the reference's
// check iterations, and at each of them the host legalization, its HPWL and
// the best snapshot (NesterovLoopState::record_metric, common.rs:485-495);
// the final binding legalizes the best snapshot (mod.rs:199).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "ep_internal.h"
#include "ep_legal.h"
#include "ep_runtime.h"

namespace ep {

#define HIPCHK(ctx, expr)                                                             \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            return fail(ctx, EP_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

int write_fixed(Ctx* c, Design& D, const double* fx, const double* fy) {
    const size_t N = (size_t)D.dev.n_mov, F = (size_t)D.dev.n_fixed;
    if (!F) return EP_OK;
    const bool f64 = D.prec == EP_PREC_FP64;
    const size_t Tb = f64 ? 8 : 4;
    std::vector<char> ax[2], xy(2 * F * Tb);
    for (int a = 0; a < 2; a++) {
        const double* src = a ? fy : fx;
        ax[a].resize(F * Tb);
        for (size_t j = 0; j < F; j++) {
            if (f64) {
                memcpy(&ax[a][j * 8], &src[j], 8);
            } else {
                const float f = (float)src[j];
                memcpy(&ax[a][j * 4], &f, 4);
            }
            memcpy(&xy[(2 * j + a) * Tb], &ax[a][j * Tb], Tb);
        }
        HIPCHK(c, hipMemcpyAsync((char*)D.dev.pos[a] + N * Tb, ax[a].data(), F * Tb, hipMemcpyHostToDevice,
                                 c->stream));
    }
    HIPCHK(c, hipMemcpyAsync((char*)D.dev.look + 2 * N * Tb, xy.data(), xy.size(), hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));  // host vectors are read until here
    return EP_OK;
}

// The solver positions (x after the step, clamped) of a resident design.
static int read_positions(Ctx* c, const Design& D, std::vector<double>& x, std::vector<double>& y) {
    const size_t N = (size_t)D.dev.n_mov;
    x.resize(N);
    y.resize(N);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int a = 0; a < 2; a++) {
        double* out = a ? y.data() : x.data();
        if (D.prec == EP_PREC_FP64) {
            HIPCHK(c, hipMemcpy(out, D.dev.pos[a], N * 8, hipMemcpyDeviceToHost));
        } else {
            std::vector<float> f(N);
            HIPCHK(c, hipMemcpy(f.data(), D.dev.pos[a], N * 4, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < N; i++) out[i] = f[i];
        }
    }
    return EP_OK;
}

static int legal_fail(Ctx* c, int rc) {
    if (rc == EP_ERR_NO_BELS) return fail(c, rc, "no free BEL of a cell's bucket (NoBelsAvailable)");
    if (rc == EP_ERR_PLACEMENT_FAILED) return fail(c, rc, "a cluster child found no free BEL (PlacementFailed)");
    if (rc == EP_ERR_NONFINITE) return fail(c, rc, "non-finite position at legalization");
    return fail(c, rc, "legalization failed");
}

}  // namespace ep

using namespace ep;

extern "C" int ep_place_legalized(ep_ctx* h, uint32_t id, const ep_config* cfg, const ep_legal_problem* lp,
                                  ep_result* res, ep_legal_result* lres, int32_t* cell_bel,
                                  int32_t* child_bel) {
    Ctx* c = (Ctx*)h;
    if (!c || !cfg || !lp || !res || !lres) return fail(c, EP_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::recursive_mutex> g(c->mu);
    if (c->server_running) return fail(c, EP_ERR_INVALID_ARG, "context is serving; stop it first");
    HIPCHK(c, hipSetDevice(c->device));
    Design* D = find_design(c, id);
    if (!D) return fail(c, EP_ERR_NOT_FOUND, "unknown design id");
    int rc = validate_config(c, cfg, D->log_cap);
    if (rc) return rc;
    const int32_t N = D->dev.n_mov, F = D->dev.n_fixed;
    const char* why = nullptr;
    if (check_legal_problem(lp, N, F, D->dev.n_nets, &why)) return fail(c, EP_ERR_INVALID_ARG, why);
    if (N > 0 && !cell_bel) return fail(c, EP_ERR_INVALID_ARG, "null cell_bel");
    if (lp->n_children > 0 && !child_bel) return fail(c, EP_ERR_INVALID_ARG, "null child_bel");
    memset(res, 0, sizeof(*res));
    lres->best_iter = -1;
    lres->legalizations = 0;
    lres->best_hpwl = std::numeric_limits<double>::infinity();
    lres->final_displacement = 0.0;
    if (N == 0) return EP_OK;  // mod.rs:80-82

    if (D->dev.n_nets > 0 && D->net_off_h.empty())
        return fail(c, EP_ERR_INVALID_ARG, "design was not uploaded with ep_design_upload");
    // the host view of the design for the legalized HPWL: the uploaded CSR,
    // the fixed locations with the cluster children's current ones
    std::vector<double> fx = D->fixed0[0], fy = D->fixed0[1];
    ep_design hd;
    memset(&hd, 0, sizeof(hd));
    hd.n_mov = N;
    hd.n_fixed = F;
    hd.n_nets = D->dev.n_nets;
    hd.n_pins = (int64_t)D->pin_cell_h.size();
    hd.net_off = D->net_off_h.data();
    hd.pin_cell = D->pin_cell_h.data();
    hd.fixed_x = D->fixed0[0].data();
    hd.fixed_y = D->fixed0[1].data();
    bool moves_fixed = false;
    for (int32_t k = 0; k < lp->n_children; k++) moves_fixed |= lp->child_fixed[k] >= 0;

    Batch* b = cached_batch(c, {D}, cfg, &rc);
    if (!b) return rc;
    if (moves_fixed && (rc = write_fixed(c, *D, fx.data(), fy.data()))) return rc;
    if ((rc = batch_reset(c, *b, cfg))) return rc;
    Legalizer L(lp->n_bels, lp->bel_x, lp->bel_y, lp->bel_bucket);
    std::vector<uint8_t> taken0((size_t)lp->n_bels, 0), taken;
    if (lp->bel_taken) taken0.assign(lp->bel_taken, lp->bel_taken + lp->n_bels);
    std::vector<int32_t> cb((size_t)N), kb((size_t)lp->n_children);
    // NesterovLoopState::new: the best positions start as the initial ones
    std::vector<double> bx, by, x, y;
    if ((rc = read_positions(c, *D, bx, by))) return rc;
    const int interval = cfg->legalize_interval > 0 ? cfg->legalize_interval : 1;
    const int max_it = cfg->max_iters;
    int last = -1;  // last iteration executed
    while (last + 1 < max_it) {
        // run to the next check iteration: 0, then every interval, and the last
        const int next = last < 0 ? 0 : std::min(last + interval, max_it - 1);
        if ((rc = batch_run_iters(c, *b, next - last, getenv("EP_NO_GRAPH") == nullptr))) return rc;
        if ((rc = read_states(c, *b))) return rc;
        const DevState& st = c->state_host[D->state_idx];
        last = st.it - 1;  // a design stopped inside the chunk skipped the rest of it
        if (st.nonfinite) break;
        if ((rc = read_positions(c, *D, x, y))) return rc;
        taken = taken0;
        double disp = 0.0;
        if ((rc = L.run(N, x.data(), y.data(), lp->cell_bucket, lp, taken.data(), cb.data(), kb.data(), &disp)))
            return legal_fail(c, rc);
        lres->legalizations++;
        // the children's new BELs: the legalized HPWL sees them, and so does the
        // WA gradient of the next iterations (collect_pin_position_xy, common.rs:268-291)
        for (int32_t k = 0; k < lp->n_children; k++)
            if (lp->child_fixed[k] >= 0) {
                fx[lp->child_fixed[k]] = lp->bel_x[kb[k]];
                fy[lp->child_fixed[k]] = lp->bel_y[kb[k]];
            }
        hd.fixed_x = fx.data();
        hd.fixed_y = fy.data();
        const double hp = legal_hpwl(&hd, lp, cb.data(), kb.data());
        if (hp < lres->best_hpwl) {  // record_metric: strictly better only
            lres->best_hpwl = hp;
            lres->best_iter = last;
            bx = x;
            by = y;
        }
        if (st.done) break;
        if (moves_fixed && (rc = write_fixed(c, *D, fx.data(), fy.data()))) return rc;
    }
    rc = batch_finalize(c, *b);
    // the design's own fixed locations back on the device
    if (moves_fixed) {
        int rc2 = write_fixed(c, *D, D->fixed0[0].data(), D->fixed0[1].data());
        if (!rc) rc = rc2;
    }
    if (rc) return rc;
    fill_result(*D, c->state_host[D->state_idx], res);
    if (res->status == EP_ERR_NONFINITE) return fail(c, res->status, "non-finite value in the placement state");
    // mod.rs:199: the final binding legalizes the best snapshot
    taken = taken0;
    if ((rc = L.run(N, bx.data(), by.data(), lp->cell_bucket, lp, taken.data(), cell_bel, child_bel,
                    &lres->final_displacement)))
        return legal_fail(c, rc);
    return EP_OK;
}
