// Per-iteration kernels of the electrostatic placer on CDNA4 (gfx950).
//
// Reference loop: crates/nextpnr/src/placer/electro_place/mod.rs:104-197.
// Each kernel covers a batch of designs; blocks find their design and their
// block of it in a block table (the host deals each design's blocks to one
// XCD, or over all of them) and skip designs whose state is done.
//
//   k_wa          WA wirelength gradient (wa.rs:51-101, common.rs:205-265):
//                 one group of G lanes (G = 2..64, a power of two) per net,
//                 pins read coalesced from padded slot arrays, max-shifted
//                 exponentials, group reductions by cross-lane shuffles; each
//                 pin's gradient goes to its own slot.  In HPWL mode the
//                 continuous HPWL (wirelength.rs:10-51) per wave.
//   k_wa_hf       the same for nets of more than 64 pins, one block per net.
//   k_wl_cell     each cell's slots summed in ascending slot order (FP64).
//   k_tile_count, k_tile_scan, k_tile_fill
//                 counting sort of the cells by 8x8-bin tile, every iteration.
//   k_scatter_pull  cell -> bin overlap density (density.rs:54-74) summed per
//                 32x32-bin super-tile in LDS, plus overflow partials
//                 (density.rs:79-86); integer bin ranges computed in FP64
//                 exactly as the oracle does.
//   k_scatter, k_dens_final  the same with global accumulation, for designs
//                 whose footprints are too wide for the super-tile halo.
//   k_gather      field gather (density.rs:171-206) fused with the penalty
//                 combine and preconditioner (mod.rs:145-156) and the
//                 Barzilai-Borwein partial sums (nesterov.rs:199-218).
//   k_ctrl        FP64 control: overflow, penalty init/growth (mod.rs:131-143),
//                 BB step (mod.rs:158-166), FISTA momentum (nesterov.rs:41-82),
//                 convergence (mod.rs:176-195), iteration log.
//   k_update      Nesterov step + clamp + next look-ahead (nesterov.rs:41-72,
//                 mod.rs:105-107,169-173).
//
// Every FAST-path sum that crosses threads is integer (2^31 fixed point) or
// in a fixed order, so a design's trajectory is the same alone and inside
// any batch.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <type_traits>

#include "ep_internal.h"

namespace ep {

template <typename T> struct Lim;
template <> struct Lim<float> { static __device__ float big() { return FLT_MAX; } };
template <> struct Lim<double> { static __device__ double big() { return DBL_MAX; } };

__device__ inline float fexp(float x) { return __expf(x); }
__device__ inline double fexp(double x) { return exp(x); }

__device__ inline int find_blk(const int32_t* pfx, int nd, int blk) {
    int lo = 0, hi = nd;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (pfx[mid] <= blk) lo = mid; else hi = mid;
    }
    return lo;
}

// The design of this block and the block's index inside the design: from
// the block's table entry (the design in the low kRefBits bits, the block
// above them; -1 for a block that only pads the XCDs to equal length), or
// (one-design views, no table) by search over the block prefix.  A block's
// work depends on this pair only, so any order of the table gives the same
// results bit for bit.
struct BlockRef {
    int d, blk;
};
__device__ inline BlockRef block_ref(const int32_t* tab, const int32_t* pfx, int nd, int p) {
    if (!tab) {
        const int d = find_blk(pfx, nd, p);
        return {d, p - pfx[d]};
    }
    const int32_t e = tab[p];
    return {e < 0 ? -1 : (int)(e & ((1 << kRefBits) - 1)), (int)(e >> kRefBits)};
}

template <typename T>
__device__ inline T shfl_xor(T v, int o) { return __shfl_xor(v, o, kWave); }

// Butterfly partner of a group reduction, by DPP where the hardware offers
// it: quad_perm for distances 1 and 2; for 4 and 8 the half-row and row
// mirrors, which pair each lane with a lane of the other half-group (the
// earlier steps left every lane of a half-group the same value, so the
// result equals the xor partner's bit for bit).  Distances 16 and 32 use
// the LDS crossbar.
template <int O>
__device__ inline int dpp_partner(int v) {
    if constexpr (O == 1) return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, true);   // quad_perm [1,0,3,2]
    else if constexpr (O == 2) return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, true);  // quad_perm [2,3,0,1]
    else if constexpr (O == 4) return __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, true);  // row_half_mirror
    else return __builtin_amdgcn_update_dpp(0, v, 0x140, 0xF, 0xF, true);                      // row_mirror
}
template <int O>
__device__ inline float partner(float v) {
    if constexpr (O <= 8) return __int_as_float(dpp_partner<O>(__float_as_int(v)));
    else return __shfl_xor(v, O, kWave);
}
template <int O>
__device__ inline double partner(double v) {
    if constexpr (O <= 8) {
        const int lo = dpp_partner<O>(__double2loint(v)), hi = dpp_partner<O>(__double2hiint(v));
        return __hiloint2double(hi, lo);
    } else {
        return __shfl_xor(v, O, kWave);
    }
}


// Deterministic block sum of NV doubles (256 threads); result in lane 0 of wave 0.
template <int NV>
__device__ inline void block_sum(double (&v)[NV], double* sh) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; q++) {
        // ascending butterfly: distances 1-8 by DPP lane moves, 16 and 32
        // through the LDS crossbar; every lane ends with the wave's sum
        double x = v[q];
        x += partner<1>(x);
        x += partner<2>(x);
        x += partner<4>(x);
        x += partner<8>(x);
        x += partner<16>(x);
        x += partner<32>(x);
        v[q] = x;
    }
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < NV; q++) sh[w * NV + q] = v[q];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; q++) v[q] = sh[q] + sh[NV + q] + sh[2 * NV + q] + sh[3 * NV + q];
    }
}

__device__ inline double sched(const DevCfg& c, double lam) {
    return lam < c.grow_thr ? lam * c.grow_mult : lam + c.grow_add;  // mod.rs:139-143
}

// ---------------------------------------------------------------- WA ----

// Gradient of one pin of a net given the net's shifted sums (exact in R to
// the reference's mean-centred formula, wa.rs:88-99):
//   d+ = e/S (1 + a (u - U/S)),  u = x - x_max, e = exp(a u)
//   d- = f/Q (1 - a (v - V/Q)),  v = x - x_min, f = exp(-a v)
// The FAST path takes 1/S and 1/Q by the hardware reciprocal (1 ulp; S and Q
// are at least 1 for a net with a pin): a correctly rounded division per sum
// cost 7 % of the WA kernel.
template <typename T>
__device__ inline T wa_pin(T a, T u, T e, T S, T U, T v, T f, T Q, T V) {
    T iS, iQ;
    if constexpr (std::is_same<T, float>::value) {
        iS = __builtin_amdgcn_rcpf(S);
        iQ = __builtin_amdgcn_rcpf(Q);
    } else {
        iS = T(1) / S;
        iQ = T(1) / Q;
    }
    return e * iS * (T(1) + a * (u - U * iS)) - f * iQ * (T(1) - a * (v - V * iQ));
}

// Two-component vector of the arithmetic type (one 8- or 16-byte access).
template <typename T> struct Vec2;
template <> struct Vec2<float> { typedef float2 type; };
template <> struct Vec2<double> { typedef double2 type; };

// Position of pin cell c: the interleaved look-ahead (WA) or the solver
// positions (HPWL); the fixed cells' locations follow the movable cells'
// entries in both, so any cell is one load without a branch.
template <typename T, bool HPWL>
__device__ inline void pin_pos(const DesignDev& D, uint32_t c, T& px, T& py) {
    if (HPWL) {
        px = ((const T*)D.pos[0])[c];
        py = ((const T*)D.pos[1])[c];
    } else {
        const typename Vec2<T>::type v = ((const typename Vec2<T>::type*)D.look)[c];
        px = v.x;
        py = v.y;
    }
}

// What a WA unit reads of its design's descriptor, loaded once at kernel
// start: loads after a store of the same kernel cannot be scalar loads.
template <typename T>
struct WaArgs {
    T a;                  // wl_coeff
    uint32_t n_mov;
    void* pg;
    const void* slot_w;
    double* part_net;
};

// Group size (pins per net) of unit w: the bucket boundaries at or below w,
// counted (the starts are non-decreasing), so the boundary loads do not wait
// on each other.
__device__ inline int wa_group(const DesignDev& D, int w) {
    int bk = 0;
#pragma unroll
    for (int q = 1; q < kMaxBuckets; q++) bk += (q < D.n_buckets && w >= D.bucket_wave_start[q]) ? 1 : 0;
    return D.bucket_group[bk];
}

// One reduction step of distance O over groups of G lanes (no-op for G <= O).
template <int O, typename T>
__device__ inline void group_step(int G, T& xmx, T& xmn, T& ymx, T& ymn) {
    if (G > O) {
        xmx = fmax(xmx, partner<O>(xmx));
        xmn = fmin(xmn, partner<O>(xmn));
        ymx = fmax(ymx, partner<O>(ymx));
        ymn = fmin(ymn, partner<O>(ymn));
    }
}
template <int O, typename T>
__device__ inline void group_sum(int G, T (&s)[8]) {
    if (G > O)
#pragma unroll
        for (int q = 0; q < 8; q++) s[q] += partner<O>(s[q]);
}

// One 64-slot unit w of the net layout (nets of G pins): the group
// reductions of its nets.  HPWL mode: h (lane 0) is the unit's partial of sum
// over nets of (max - min) x + y of the solver positions.  Else (gx, gy) is
// the WA gradient of the lane's pin, to be stored to the pin's own slot when
// the lane holds a movable pin; each cell later sums its slots in a fixed
// order (cell -> slot CSR), so there are no atomics and the sum is
// deterministic.
template <typename T, bool HPWL>
__device__ inline void wa_unit(const WaArgs<T>& A, int G, int lane, uint32_t c, T px, T py, double& h,
                               T& gx, T& gy) {
    const bool valid = c != kPad;
    const T big = Lim<T>::big();
    T xmx = valid ? px : -big, xmn = valid ? px : big;
    T ymx = valid ? py : -big, ymn = valid ? py : big;
    group_step<1>(G, xmx, xmn, ymx, ymn);
    group_step<2>(G, xmx, xmn, ymx, ymn);
    group_step<4>(G, xmx, xmn, ymx, ymn);
    group_step<8>(G, xmx, xmn, ymx, ymn);
    group_step<16>(G, xmx, xmn, ymx, ymn);
    group_step<32>(G, xmx, xmn, ymx, ymn);
    if (HPWL) {
        // wave padding forms whole groups without a valid pin (max < min): skip them
        const bool has = xmx >= xmn;
        h = ((lane & (G - 1)) == 0 && has) ? ((double)xmx - (double)xmn) + ((double)ymx - (double)ymn) : 0.0;
        for (int o = 32; o >= 1; o >>= 1) h += __shfl_xor(h, o, kWave);
        return;
    }
    const T a = A.a;
    const T ux = px - xmx, vx = px - xmn, uy = py - ymx, vy = py - ymn;
    const T ex = valid ? fexp(a * ux) : T(0), fx = valid ? fexp(-a * vx) : T(0);
    const T ey = valid ? fexp(a * uy) : T(0), fy = valid ? fexp(-a * vy) : T(0);
    T s[8] = {ex, ux * ex, fx, vx * fx, ey, uy * ey, fy, vy * fy};
    group_sum<1>(G, s);
    group_sum<2>(G, s);
    group_sum<4>(G, s);
    group_sum<8>(G, s);
    group_sum<16>(G, s);
    group_sum<32>(G, s);
    gx = wa_pin<T>(a, ux, ex, s[0], s[1], vx, fx, s[2], s[3]);
    gy = wa_pin<T>(a, uy, ey, s[4], s[5], vy, fy, s[6], s[7]);
}

// Each wave runs kWaUnits consecutive units: the slot loads of all of them,
// then their position loads, are issued before any unit's reductions, and
// every store follows all of the reductions, so a wave keeps kWaUnits
// dependent load chains in flight and never waits on its own stores.
template <typename T, bool FIX, bool HPWL>
__global__ __launch_bounds__(256) void k_wa(BatchDev b, int final_pass) {
    const BlockRef ref = block_ref(b.blk_wa, b.pfx_wa, b.nd, blockIdx.x);
    if (ref.d < 0) return;
    const int d = ref.d;
    const DesignDev& D = b.d[d];
    const DevState* st = D.st;
    if (HPWL) {
        if (!final_pass && !(st->run_update && st->check && D.cfg.log_hpwl)) return;
    } else if (st->done) {
        return;
    }
    // the wave index through readfirstlane: the unit, its bucket and group
    // size are then scalars and the descriptor fields they select scalar loads
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int w0 = (ref.blk * 4 + wv) * kWaUnits;
    if (w0 >= D.n_wa_waves) return;  // uniform per wave; no block barrier below
    const int lane = threadIdx.x & 63;
    const int nu = min(kWaUnits, D.n_wa_waves - w0);
    const WaArgs<T> A{(T)D.cfg.wl_coeff, (uint32_t)D.n_mov, D.pg, D.slot_w, D.part_net};
    uint32_t c[kWaUnits];
    int G[kWaUnits];
    T px[kWaUnits], py[kWaUnits];
#pragma unroll
    for (int u = 0; u < kWaUnits; u++) {
        c[u] = u < nu ? D.slots[(size_t)(w0 + u) * kWave + lane] : kPad;
        G[u] = wa_group(D, w0 + u);
    }
    // a padding slot loads cell 0's entry (value unused)
#pragma unroll
    for (int u = 0; u < kWaUnits; u++) pin_pos<T, HPWL>(D, c[u] != kPad ? c[u] : 0u, px[u], py[u]);
    double h[kWaUnits];
    T gx[kWaUnits], gy[kWaUnits];
#pragma unroll
    for (int u = 0; u < kWaUnits; u++)
        if (u < nu) wa_unit<T, HPWL>(A, G[u], lane, c[u], px[u], py[u], h[u], gx[u], gy[u]);
#pragma unroll
    for (int u = 0; u < kWaUnits; u++) {
        if (u >= nu) break;
        const int w = w0 + u;
        if (HPWL) {
            if (lane == 0) A.part_net[w] = h[u];
        } else if (c[u] != kPad && c[u] < A.n_mov) {
            const uint32_t slot = (uint32_t)w * kWave + lane;
            T wt = T(1);
            if (A.slot_w) wt = ((const T*)A.slot_w)[slot];  // net weight, common.rs:246-261 (absent: 1)
            typename Vec2<T>::type v;
            v.x = A.slot_w ? wt * gx[u] : gx[u];
            v.y = A.slot_w ? wt * gy[u] : gy[u];
            ((typename Vec2<T>::type*)A.pg)[slot] = v;
        }
    }
}

// Nets of more than 64 pins: one block of kHfThreads per net, two passes
// over its pins (extremes, then exponential sums), and the net's statistics
// written for the cells to form their pins' gradients from.  A net of up to
// kHfThreads * kHfCache pins keeps each thread's pin positions in registers
// after the first pass; larger nets read them again.  The block reductions
// run over the waves in a fixed order.
constexpr int kHfThreads = 512;
constexpr int kHfWaves = kHfThreads / kWave;
constexpr int kHfCache = 10;

template <int NV>
__device__ inline void hf_block_sum(double (&v)[NV], double* sh, double* out) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; q++) {
        double x = v[q];
        for (int o = 32; o >= 1; o >>= 1) x += __shfl_xor(x, o, kWave);
        if (lane == 0) sh[w * NV + q] = x;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double t = 0.0;
        for (int k = 0; k < kHfWaves; k++) t += sh[k * NV + threadIdx.x];
        out[threadIdx.x] = t;
    }
    __syncthreads();
}

template <typename T, bool FIX, bool HPWL>
__global__ __launch_bounds__(kHfThreads) void k_wa_hf(BatchDev b, int final_pass) {
    __shared__ double sh[kHfWaves * 8];
    __shared__ double tot[8];
    const BlockRef ref = block_ref(b.blk_hf, b.pfx_hf, b.nd, blockIdx.x);
    if (ref.d < 0) return;
    const int d = ref.d;
    const DesignDev& D = b.d[d];
    const DevState* st = D.st;
    if (HPWL) {
        if (!final_pass && !(st->run_update && st->check && D.cfg.log_hpwl)) return;
    } else if (st->done) {
        return;
    }
    const int e = ref.blk;
    if (e >= D.n_hf) return;
    const uint32_t p0 = D.hf_off[e], p1 = D.hf_off[e + 1];
    const bool cached = p1 - p0 <= (uint32_t)(kHfThreads * kHfCache);
    const T big = Lim<T>::big();
    T cx[kHfCache], cy[kHfCache];
    double mm[4] = {-(double)big, -(double)big, -(double)big, -(double)big};  // xmax, -xmin, ymax, -ymin
    auto extremes = [&](T px, T py) {
        mm[0] = fmax(mm[0], (double)px);
        mm[1] = fmax(mm[1], -(double)px);
        mm[2] = fmax(mm[2], (double)py);
        mm[3] = fmax(mm[3], -(double)py);
    };
    if (cached) {
#pragma unroll
        for (int q = 0; q < kHfCache; q++) {
            const uint32_t p = p0 + threadIdx.x + (uint32_t)q * kHfThreads;
            cx[q] = cy[q] = T(0);
            if (p < p1) {
                pin_pos<T, HPWL>(D, D.hf_pins[p], cx[q], cy[q]);
                extremes(cx[q], cy[q]);
            }
        }
    } else {
        for (uint32_t p = p0 + threadIdx.x; p < p1; p += kHfThreads) {
            T px, py;
            pin_pos<T, HPWL>(D, D.hf_pins[p], px, py);
            extremes(px, py);
        }
    }
    {  // block max, fixed order
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            double x = mm[q];
            for (int o = 32; o >= 1; o >>= 1) x = fmax(x, __shfl_xor(x, o, kWave));
            if (lane == 0) sh[w * 4 + q] = x;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; q++) {
            double x = sh[q];
            for (int k = 1; k < kHfWaves; k++) x = fmax(x, sh[k * 4 + q]);
            mm[q] = x;
        }
        __syncthreads();
    }
    const T xmx = (T)mm[0], xmn = (T)(-mm[1]), ymx = (T)mm[2], ymn = (T)(-mm[3]);
    if (HPWL) {
        if (threadIdx.x == 0)
            D.part_net[D.n_wa_waves + e] = ((double)xmx - (double)xmn) + ((double)ymx - (double)ymn);
        return;
    }
    const T a = (T)D.cfg.wl_coeff;
    double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    auto sums = [&](T px, T py) {
        T ux = px - xmx, vx = px - xmn, uy = py - ymx, vy = py - ymn;
        T ex = fexp(a * ux), fx = fexp(-a * vx), ey = fexp(a * uy), fy = fexp(-a * vy);
        s[0] += ex; s[1] += ux * ex; s[2] += fx; s[3] += vx * fx;
        s[4] += ey; s[5] += uy * ey; s[6] += fy; s[7] += vy * fy;
    };
    if (cached) {
#pragma unroll
        for (int q = 0; q < kHfCache; q++)
            if (p0 + threadIdx.x + (uint32_t)q * kHfThreads < p1) sums(cx[q], cy[q]);
    } else {
        for (uint32_t p = p0 + threadIdx.x; p < p1; p += kHfThreads) {
            T px, py;
            pin_pos<T, HPWL>(D, D.hf_pins[p], px, py);
            sums(px, py);
        }
    }
    hf_block_sum<8>(s, sh, tot);
    // the net's extremes, sums and weight; each cell derives its pins'
    // gradients from them (gwl_of), so no per-pin slot is written
    if (threadIdx.x == 0) {
        T* out = (T*)D.hf_stat + (size_t)e * kHfStat;
        out[0] = xmx; out[1] = xmn; out[2] = ymx; out[3] = ymn;
#pragma unroll
        for (int q = 0; q < 8; q++) out[4 + q] = (T)tot[q];
        out[12] = D.slot_w ? ((const T*)D.slot_w)[(size_t)D.n_wa_waves * kWave + p0] : T(1);
    }
}

// ----------------------------------------------------------- density ----

struct BinBox {
    int bx0, bx1, by0, by1;
    double xl, xh, yl, yh;
};

// density.rs:20-37 with the area / bin-size extension; FP64 so the integer
// ranges equal the oracle's bit for bit for the same input positions.
__device__ inline BinBox bin_box(const DesignDev& D, double cx, double cy, double hs) {
    BinBox r;
    r.xl = (cx - hs) * D.rx;
    r.xh = (cx + hs) * D.rx;
    r.yl = (cy - hs) * D.ry;
    r.yh = (cy + hs) * D.ry;
    int b = (int)floor(r.xl);
    r.bx0 = b < 0 ? 0 : b;
    double c = ceil(r.xh);
    r.bx1 = c <= 0.0 ? 0 : (c >= (double)D.Bx ? D.Bx : (int)c);
    b = (int)floor(r.yl);
    r.by0 = b < 0 ? 0 : b;
    c = ceil(r.yh);
    r.by1 = c <= 0.0 ? 0 : (c >= (double)D.By ? D.By : (int)c);
    return r;
}

// Density accumulators.  FAST: 64-bit fixed point 2^-31 (one word per bin);
// a contribution is one bin's overlap, at most 1, so it converts through 32
// bits (one rounding and one conversion instruction where a 64-bit
// conversion takes a sequence).  FP64: exact 128-bit fixed point, an integer
// word and a 2^-64 fraction word per bin (FIX = false); the fraction's carry
// goes to the integer word, so the sum is the exact sum of the truncated
// contributions in any order and the plain placer's density is deterministic
// to the last bit.
template <bool FIX>
__device__ inline void acc_add(unsigned long long* whole, unsigned long long* frac, double w) {
    if constexpr (FIX) {
        atomicAdd(whole, (unsigned long long)(uint32_t)rint(w * kFixScale));
    } else {
        const double ip = floor(w);
        const unsigned long long f = __double2ull_rz((w - ip) * 18446744073709551616.0);
        const unsigned long long old = atomicAdd(frac, f);
        const unsigned long long carry = (unsigned long long)ip + ((old + f < old) ? 1ull : 0ull);
        if (carry) atomicAdd(whole, carry);
    }
}
template <bool FIX>
__device__ inline double acc_value(unsigned long long whole, unsigned long long frac) {
    if constexpr (FIX) return (double)(long long)whole * (1.0 / kFixScale);
    else return (double)whole + (double)frac * 5.421010862427522e-20;  // 2^-64
}

// density.rs:41-48, as the product of the overlap along x with bin column
// bx and along y with bin row by: the gather forms a cell's three column
// widths and row heights once and multiplies them per bin (the same bits as
// ov_area).  The pull scatter keeps ov_area per bin: there the per-axis form
// measured slower.
__device__ inline double ov_x(const BinBox& r, int bx) {
    const double o = fmin(r.xh, (double)bx + 1.0) - fmax(r.xl, (double)bx);
    return o < 0.0 ? 0.0 : o;
}
__device__ inline double ov_y(const BinBox& r, int by) {
    const double o = fmin(r.yh, (double)by + 1.0) - fmax(r.yl, (double)by);
    return o < 0.0 ? 0.0 : o;
}
__device__ inline double ov_area(const BinBox& r, int bx, int by) { return ov_x(r, bx) * ov_y(r, by); }

// Scatter for designs whose widest footprint is too wide for the pull
// scatter below.  One block covers kScatCells consecutive cells of this
// iteration's tile order `perm`, so its cells sit in a compact group of
// bins; when the block's bin bounding box fits
// kScatBins the contributions are summed in LDS and flushed with one global
// atomic per touched bin, otherwise each contribution goes to global memory
// directly.  All sums are fixed point (FAST), so neither the order nor the
// path changes the result.
template <typename T, bool FIX>
__global__ __launch_bounds__(256) void k_scatter(BatchDev b) {
    typedef unsigned long long Acc;
    // FAST: kScatBins one-word bins; FP64: half as many two-word bins
    constexpr int kCap = FIX ? kScatBins : kScatBins / 2;
    __shared__ Acc win[kScatBins];
    __shared__ int box[4];
    const BlockRef ref = block_ref(b.blk_scat, b.pfx_scat, b.nd, blockIdx.x);
    if (ref.d < 0) return;
    const int d = ref.d;
    const DesignDev& D = b.d[d];
    if (D.st->done) return;
    const int c0 = ref.blk * kScatCells;
    const uint32_t* PERM = D.perm[(D.st->it + 1) & 1];
    typedef typename Vec2<T>::type V;
    const V* LXY = (const V*)D.look;
    const size_t nb = (size_t)D.Bx * D.By;
    Acc* GW = (Acc*)D.dacc;       // integer / fixed-point words
    Acc* GF = GW + nb;            // FP64 fraction words
    if (threadIdx.x == 0) {
        box[0] = box[2] = 0x7fffffff;
        box[1] = box[3] = -1;
    }
    __syncthreads();
    int lo_x = 0x7fffffff, hi_x = -1, lo_y = 0x7fffffff, hi_y = -1;
    for (int k = threadIdx.x; k < kScatCells; k += blockDim.x) {
        if (c0 + k >= D.n_mov) break;
        const int i = PERM[c0 + k];
        const V l = LXY[i];
        const BinBox r = bin_box(D, (double)l.x, (double)l.y, (double)D.hs[i]);
        if (r.bx1 > r.bx0 && r.by1 > r.by0) {
            lo_x = min(lo_x, r.bx0); hi_x = max(hi_x, r.bx1);
            lo_y = min(lo_y, r.by0); hi_y = max(hi_y, r.by1);
        }
    }
    atomicMin(&box[0], lo_x); atomicMax(&box[1], hi_x);
    atomicMin(&box[2], lo_y); atomicMax(&box[3], hi_y);
    __syncthreads();
    const int bx0 = box[0], by0 = box[2];
    const int ww = box[1] - box[0], hh = box[3] - box[2];
    const bool local = ww > 0 && hh > 0 && (long long)ww * hh <= kCap;
    if (local)
        for (int k = threadIdx.x; k < (FIX ? 1 : 2) * ww * hh; k += blockDim.x) win[k] = Acc(0);
    __syncthreads();
    for (int k = threadIdx.x; k < kScatCells; k += blockDim.x) {
        if (c0 + k >= D.n_mov) break;
        const int i = PERM[c0 + k];
        const V l = LXY[i];
        const BinBox r = bin_box(D, (double)l.x, (double)l.y, (double)D.hs[i]);
        for (int by = r.by0; by < r.by1; by++)
            for (int bx = r.bx0; bx < r.bx1; bx++) {
                const double w = ov_area(r, bx, by);
                if (local) {
                    const int q = (by - by0) * ww + (bx - bx0);
                    acc_add<FIX>(&win[q], &win[ww * hh + q], w);
                } else {
                    const size_t q = (size_t)by * D.Bx + bx;
                    acc_add<FIX>(&GW[q], &GF[q], w);
                }
            }
    }
    if (!local) return;
    __syncthreads();
    for (int k = threadIdx.x; k < ww * hh; k += blockDim.x) {
        const size_t q = (size_t)(by0 + k / ww) * D.Bx + bx0 + k % ww;
        const Acc v = win[k];
        if (FIX) {
            if (v != Acc(0)) atomicAdd(&GW[q], v);
        } else {
            // the window's fraction word goes in with its carry, the whole word after
            const Acc f = win[ww * hh + k];
            Acc carry = v;
            if (f != Acc(0)) {
                const Acc old = atomicAdd(&GF[q], f);
                carry += (old + f < old) ? 1ull : 0ull;
            }
            if (carry) atomicAdd(&GW[q], carry);
        }
    }
}

template <typename T, bool FIX>
__global__ __launch_bounds__(256) void k_dens_final(BatchDev b) {
    __shared__ double sh[4 * kBinParts];
    const BlockRef ref = block_ref(b.blk_bin, b.pfx_bin, b.nd, blockIdx.x);
    if (ref.d < 0) return;
    const int d = ref.d;
    const DesignDev& D = b.d[d];
    if (D.st->done) return;
    const int blk = ref.blk;
    const int i = blk * kBlock + threadIdx.x;
    const int nb = D.Bx * D.By;
    double v = 0.0;
    if (i < nb) {
        unsigned long long* a = (unsigned long long*)D.dacc;
        if constexpr (FIX) {
            v = acc_value<true>(a[i], 0ull);
        } else {
            v = acc_value<false>(a[i], a[(size_t)nb + i]);
            a[(size_t)nb + i] = 0ull;
        }
        a[i] = 0ull;
        ((T*)D.dens)[i] = (T)v;
    }
    double p[kBinParts] = {v, v > 1.0 ? v - 1.0 : 0.0};
    block_sum<kBinParts>(p, sh);
    if (threadIdx.x == 0) {
        D.part_bin[(size_t)blk * kBinParts + 0] = p[0];
        D.part_bin[(size_t)blk * kBinParts + 1] = p[1];
    }
}

// ------------------------------------------------------- tile sort -----

// Morton interleave of two 12-bit tile coordinates.
__device__ inline uint32_t morton2(uint32_t x, uint32_t y) {
    x = (x | (x << 8)) & 0x00FF00FFu; x = (x | (x << 4)) & 0x0F0F0F0Fu;
    x = (x | (x << 2)) & 0x33333333u; x = (x | (x << 1)) & 0x55555555u;
    y = (y | (y << 8)) & 0x00FF00FFu; y = (y | (y << 4)) & 0x0F0F0F0Fu;
    y = (y | (y << 2)) & 0x33333333u; y = (y | (y << 1)) & 0x55555555u;
    return x | (y << 1);
}

template <typename T>
__device__ inline uint32_t tile_key(const DesignDev& D, T x, T y, float hs) {
    const BinBox r = bin_box(D, (double)x, (double)y, (double)hs);
    const int bx = min(r.bx0, D.Bx - 1), by = min(r.by0, D.By - 1);
    const uint32_t k = morton2((uint32_t)bx >> 3, (uint32_t)by >> 3);
    return k < (uint32_t)D.n_keys ? k : (uint32_t)D.n_keys - 1;
}

// One atomic per distinct key of a wave: lanes with equal keys are counted
// together (the previous order keeps them mostly equal), and each lane
// receives its own slot (the key's count before the add + its rank among the
// lanes of its key).
__device__ inline uint32_t wave_agg_add(uint32_t* ctr, uint32_t key, bool valid) {
    const int lane = threadIdx.x & 63;
    unsigned long long act = __ballot(valid);
    // group the lanes by key (register work only), then every group's leader
    // adds its group's count in one atomic instruction: one round trip for
    // the wave however many keys it holds
    bool leader = false;
    uint32_t cnt = 0, rank = 0;
    int lead = 0;
    while (act) {
        const int l = __ffsll((long long)act) - 1;
        const uint32_t lk = __shfl(key, l, kWave);
        const unsigned long long same = __ballot(valid && key == lk) & act;
        if (lane == l) {
            leader = true;
            cnt = (uint32_t)__popcll(same);
        }
        if ((same >> lane) & 1ull) {
            lead = l;
            rank = (uint32_t)__popcll(same & ((1ull << lane) - 1ull));
        }
        act &= ~same;
    }
    uint32_t base = 0;
    if (leader) base = atomicAdd(&ctr[key], cnt);
    return __shfl(base, lead, kWave) + rank;
}

// Counting sort of the cells by tile key, every iteration: count, scan, fill.
// The count pass walks the cells in the previous iteration's order, in which
// neighbours mostly share a tile, so a wave sees few distinct keys; it
// gathers each cell's position once (the half sides are kept in the order of
// each permutation by the fill that wrote it), stages key, position and the
// cell's offset inside its tile by position, so the fill pass reads only
// contiguous arrays and places each cell at tstart[key] + offset.
template <typename T>
__global__ __launch_bounds__(256) void k_tile_count(BatchDev b) {
    const BlockRef ref = block_ref(b.blk_cell, b.pfx_cell, b.nd, blockIdx.x);
    if (ref.d < 0) return;
    const int d = ref.d;
    const DesignDev& D = b.d[d];
    if (D.st->done) return;
    const int k = ref.blk * kBlock + threadIdx.x;
    const bool valid = k < D.n_mov;
    uint32_t key = 0u;
    if (valid) {
        const int par = D.st->it & 1;
        const uint32_t i = D.perm[par][k];
        const typename Vec2<T>::type l = ((const typename Vec2<T>::type*)D.look)[i];
        const T x = l.x, y = l.y;
        const float hs = D.sorted_hs[par][k];  // hs[i], already in this order
        key = tile_key<T>(D, x, y, hs);
        D.keys[k] = key;
        T* sxy = (T*)D.stage_xy;
        sxy[2 * (size_t)k] = x;
        sxy[2 * (size_t)k + 1] = y;
        D.stage_hs[k] = hs;
    }
    // the counter's old value is this cell's offset inside its tile, so
    // the fill pass needs no atomics of its own
    const uint32_t ofs = wave_agg_add(D.tcount, key, valid);
    if (valid) D.tofs[k] = ofs;
}

// One block per design: tstart = exclusive scan of the counts (with the
// total at n_keys), counts cleared for the next sort.
// Each thread owns a contiguous run of keys (16-byte vectors when the key
// count is a multiple of 1,024, as every power of four from 1,024 up is);
// the run totals are scanned with wave shuffles and one LDS step across the
// four waves.  The two arrays are distinct buffers (restrict), so a run's
// loads are all issued before its stores.
__global__ __launch_bounds__(256) void k_tile_scan(BatchDev b) {
    __shared__ uint32_t wsum[4];
    const DesignDev& D = b.d[blockIdx.x];
    if (D.st->done) return;
    uint32_t* __restrict__ cnt = D.tcount;
    uint32_t* __restrict__ ts = D.tstart;
    const int n = D.n_keys;
    const bool vec = (n & 1023) == 0;
    const int per = (n + 255) / 256;
    const int lo = min((int)threadIdx.x * per, n), hi = min(lo + per, n);
    uint32_t sum = 0;
    if (vec) {
        for (int k = lo; k < hi; k += 4) {
            const uint4 v = *(const uint4*)&cnt[k];
            sum += (v.x + v.y) + (v.z + v.w);
        }
    } else {
        for (int k = lo; k < hi; k++) sum += cnt[k];
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t v = sum;  // inclusive scan of the run totals within the wave
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(v, o, kWave);
        if (lane >= o) v += t;
    }
    if (lane == 63) wsum[w] = v;
    __syncthreads();
    uint32_t base = 0;
    for (int q = 0; q < w; q++) base += wsum[q];
    if (threadIdx.x == 255) ts[n] = base + v;
    uint32_t run = base + v - sum;
    if (vec) {
        for (int k = lo; k < hi; k += 4) {
            const uint4 c = *(const uint4*)&cnt[k];
            uint4 st;
            st.x = run; run += c.x;
            st.y = run; run += c.y;
            st.z = run; run += c.z;
            st.w = run; run += c.w;
            *(uint4*)&ts[k] = st;
            *(uint4*)&cnt[k] = make_uint4(0u, 0u, 0u, 0u);
        }
    } else {
        for (int k = lo; k < hi; k++) {
            const uint32_t c = cnt[k];
            ts[k] = run;
            cnt[k] = 0u;
            run += c;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_tile_fill(BatchDev b) {
    const BlockRef ref = block_ref(b.blk_cell, b.pfx_cell, b.nd, blockIdx.x);
    if (ref.d < 0) return;
    const int d = ref.d;
    const DesignDev& D = b.d[d];
    if (D.st->done) return;
    const int k = ref.blk * kBlock + threadIdx.x;
    const bool valid = k < D.n_mov;
    const int par = D.st->it & 1;
    const uint32_t i = valid ? D.perm[par][k] : 0u;
    // the staged copy is loaded beside the key, not after the position
    typedef typename Vec2<T>::type V;
    const V xy = valid ? ((const V*)D.stage_xy)[k] : V{};
    const float hs = valid ? D.stage_hs[k] : 0.0f;
    const uint32_t pos = valid ? D.tstart[D.keys[k]] + D.tofs[k] : 0u;
    if (valid && pos < (uint32_t)D.n_mov) {
        D.perm[par ^ 1][pos] = i;
        ((V*)D.sorted_xy)[pos] = xy;
        D.sorted_hs[par ^ 1][pos] = hs;
    }
}

// Pull scatter: one block per 32x32-bin super-tile (4x4 sort tiles, one
// contiguous Morton key range).  Every cell whose footprint reaches the
// super-tile has its lower-left bin in it or in the one-tile halo to its
// left and below (footprints span at most 9 bins), so the block reads the
// super-tile's key range plus the 9 halo tiles of the sorted copies, sums
// the overlaps that fall inside its own 1,024 bins in LDS (fixed point on the
// FAST path, hence independent of order), and writes the density and its
// overflow partials with plain stores: no global atomics.
template <typename T, bool FIX>
__global__ __launch_bounds__(256) void k_scatter_pull(BatchDev b) {
    typedef unsigned long long Acc;
    // FAST: one fixed-point word per bin; FP64: the whole words, then the fractions
    __shared__ Acc win[(FIX ? 1 : 2) * 32 * 32];
    __shared__ double sh[4 * kBinParts];
    const BlockRef ref = block_ref(b.blk_tile, b.pfx_tile, b.nd, blockIdx.x);
    if (ref.d < 0) return;
    const int d = ref.d;
    const DesignDev& D = b.d[d];
    if (D.st->done) return;
    const int t = ref.blk;
    const int stx = (D.tiles_x + 3) / 4;
    const int sx = t % stx, sy = t / stx;
    const int x0 = sx * 32, y0 = sy * 32;
    // the ten key ranges: q = 0 the 4x4 block of tiles (Morton keys 16*m ..
    // 16*m+15), q = 1..9 the halo tiles (left column and bottom row, corner
    // included).  Their bounds are loaded at once, by ten threads, and the
    // cells of all ten are then walked as one index space, so every cell
    // load of the block is issued in the same pass.
    __shared__ uint32_t rk0[10], rlen[10];
    if (threadIdx.x < 10) {
        const int q = threadIdx.x;
        uint32_t k0 = 0, k1 = 0;
        if (q == 0) {
            const uint32_t m = morton2((uint32_t)sx, (uint32_t)sy) * 16u;
            if (m < (uint32_t)D.n_keys) {
                k0 = D.tstart[m];
                k1 = D.tstart[min(m + 16u, (uint32_t)D.n_keys)];
            }
        } else {
            int hx, hy;
            if (q <= 5) { hx = sx * 4 - 1; hy = sy * 4 - 1 + (q - 1); }  // left column, from the corner
            else { hx = sx * 4 + (q - 6); hy = sy * 4 - 1; }            // bottom row
            if (hx >= 0 && hy >= 0) {
                const uint32_t key = morton2((uint32_t)hx, (uint32_t)hy);
                if (key < (uint32_t)D.n_keys) {
                    k0 = D.tstart[key];
                    k1 = D.tstart[key + 1];
                }
            }
        }
        rk0[q] = k0;
        rlen[q] = k1 - k0;
    }
    for (int k = threadIdx.x; k < (FIX ? 1 : 2) * 32 * 32; k += blockDim.x) win[k] = Acc(0);
    __syncthreads();
    uint32_t pre[11];  // start of range q in the joint index space
    pre[0] = 0;
#pragma unroll
    for (int q = 0; q < 10; q++) pre[q + 1] = pre[q] + rlen[q];
    const T* SXY = (const T*)D.sorted_xy;
    const float* SHS = D.sorted_hs[(D.st->it & 1) ^ 1];
    for (uint32_t j = threadIdx.x; j < pre[10]; j += blockDim.x) {
        uint32_t k = 0;
#pragma unroll
        for (int q = 0; q < 10; q++)  // the last range starting at or before j holds it
            if (j >= pre[q]) k = rk0[q] + (j - pre[q]);
        const BinBox r = bin_box(D, (double)SXY[2 * (size_t)k], (double)SXY[2 * (size_t)k + 1],
                                 (double)SHS[k]);
        const int bx0 = max(r.bx0, x0), bx1 = min(r.bx1, x0 + 32);
        const int by0 = max(r.by0, y0), by1 = min(r.by1, y0 + 32);
        for (int by = by0; by < by1; by++)
            for (int bx = bx0; bx < bx1; bx++) {
                const double w = ov_area(r, bx, by);
                const int q = (by - y0) * 32 + (bx - x0);
                acc_add<FIX>(&win[q], &win[FIX ? q : 32 * 32 + q], w);
            }
    }
    __syncthreads();
    double p[kBinParts] = {0.0, 0.0};
    for (int k = threadIdx.x; k < 32 * 32; k += blockDim.x) {
        const int bx = x0 + (k & 31), by = y0 + (k >> 5);
        if (bx < D.Bx && by < D.By) {
            const double v = acc_value<FIX>(win[k], FIX ? 0ull : win[32 * 32 + k]);
            ((T*)D.dens)[(size_t)by * D.Bx + bx] = (T)v;
            p[0] += v;
            p[1] += v > 1.0 ? v - 1.0 : 0.0;
        }
    }
    block_sum<kBinParts>(p, sh);
    if (threadIdx.x == 0) {
        D.part_bin[(size_t)t * kBinParts + 0] = p[0];
        D.part_bin[(size_t)t * kBinParts + 1] = p[1];
    }
}

// ------------------------------------------------------------ gather ----

// WA gradient of cell i: its pin slots (nets of <= 64 pins) summed in
// ascending slot order, then its pins on high-fanout nets, in ascending net
// order, each formed from the net's statistics exactly as a slot would hold
// it.  The slot indices and slot values of up to eight pins (then four) are
// loaded before any is added, so their dependent load pairs are in flight at
// once; the additions keep the sequential order.
template <typename T>
__device__ inline void gwl_of(const DesignDev& D, int i, double& gx, double& gy) {
    typedef typename Vec2<T>::type V;
    const V* pg = (const V*)D.pg;
    double sx = 0.0, sy = 0.0;
    // both CSR bounds in one round of loads (the high-fanout bounds are not
    // left to wait behind the slot sums)
    uint32_t k = D.cs_off[i];
    const uint32_t e = D.cs_off[i + 1];
    const uint32_t h0 = D.chf_off[i], h1 = D.chf_off[i + 1];
    {
        // the first kCellSlots slots (all but a few cells in a thousand have
        // no more; with eight, nearly every wave held a cell with more) come
        // from the fixed-stride copy, loaded with the bounds: their values
        // are the second round of loads.  Past the cell's last pin the row
        // names the zero slot, whose +0.0 leaves the sum unchanged bit for
        // bit (it starts at +0.0 and never becomes -0.0).
        static_assert(kCellSlots % 4 == 0, "16-byte loads of the fixed-stride row");
        constexpr int kQ = kCellSlots / 4;
        uint32_t sl[kCellSlots];
#pragma unroll
        for (int q = 0; q < kQ; q++) {
            const uint4 a = ((const uint4*)D.cs8)[kQ * (size_t)i + q];
            sl[4 * q] = a.x;
            sl[4 * q + 1] = a.y;
            sl[4 * q + 2] = a.z;
            sl[4 * q + 3] = a.w;
        }
        V vs[kCellSlots];
#pragma unroll
        for (int q = 0; q < kCellSlots; q++) vs[q] = pg[sl[q]];
#pragma unroll
        for (int q = 0; q < kCellSlots; q++) {
            sx += (double)vs[q].x;
            sy += (double)vs[q].y;
        }
        k += min(e - k, (uint32_t)kCellSlots);
    }
    for (; k + 4 <= e; k += 4) {
        const uint32_t s0 = D.cs_idx[k], s1 = D.cs_idx[k + 1];
        const uint32_t s2 = D.cs_idx[k + 2], s3 = D.cs_idx[k + 3];
        const V v0 = pg[s0], v1 = pg[s1], v2 = pg[s2], v3 = pg[s3];
        sx += (double)v0.x; sy += (double)v0.y;
        sx += (double)v1.x; sy += (double)v1.y;
        sx += (double)v2.x; sy += (double)v2.y;
        sx += (double)v3.x; sy += (double)v3.y;
    }
    const int rest = (int)(e - k);
    if (rest > 0) {
        const uint32_t s0 = D.cs_idx[k];
        const uint32_t s1 = rest > 1 ? D.cs_idx[k + 1] : s0;
        const uint32_t s2 = rest > 2 ? D.cs_idx[k + 2] : s0;
        const V v0 = pg[s0], v1 = pg[s1], v2 = pg[s2];
        sx += (double)v0.x; sy += (double)v0.y;
        if (rest > 1) { sx += (double)v1.x; sy += (double)v1.y; }
        if (rest > 2) { sx += (double)v2.x; sy += (double)v2.y; }
    }
    if (h1 > h0) {
        const V l = ((const V*)D.look)[i];
        const T a = (T)D.cfg.wl_coeff;
        for (uint32_t k = h0; k < h1; k++) {
            const T* st = (const T*)D.hf_stat + (size_t)D.chf_net[k] * kHfStat;
            const T ux = l.x - st[0], vx = l.x - st[1], uy = l.y - st[2], vy = l.y - st[3];
            const T ex = fexp(a * ux), fx = fexp(-a * vx), ey = fexp(a * uy), fy = fexp(-a * vy);
            T gxp = wa_pin<T>(a, ux, ex, st[4], st[5], vx, fx, st[6], st[7]);
            T gyp = wa_pin<T>(a, uy, ey, st[8], st[9], vy, fy, st[10], st[11]);
            if (D.slot_w) {  // net weight, common.rs:246-261
                gxp = st[12] * gxp;
                gyp = st[12] * gyp;
            }
            sx += (double)gxp;
            sy += (double)gyp;
        }
    }
    gx = sx;
    gy = sy;
}

// The WA gradient of every cell (its slots summed, gwl_of), written per cell
// so the gather reads one contiguous pair; launched right after the WA
// kernels, beside the density chain.
template <typename T>
__global__ __launch_bounds__(256) void k_wl_cell(BatchDev b) {
    const BlockRef ref = block_ref(b.blk_cell, b.pfx_cell, b.nd, blockIdx.x);
    if (ref.d < 0) return;
    const int d = ref.d;
    const DesignDev& D = b.d[d];
    if (D.st->done) return;
    const int i = ref.blk * kBlock + threadIdx.x;
    if (i >= D.n_mov) return;
    double2 v;
    gwl_of<T>(D, i, v.x, v.y);
    typename Vec2<T>::type w;  // summed in FP64, kept in the arithmetic type
    w.x = (T)v.x;
    w.y = (T)v.y;
    ((typename Vec2<T>::type*)D.gwl)[i] = w;
}

// MODE 0: gather; MODE 1: fix (penalty set by ctrl this iteration);
// MODE 2: evaluation export of the density gradient only.
template <typename T, bool FIX, int MODE>
__global__ __launch_bounds__(256) void k_gather(BatchDev b, double* ex_out, double* ey_out) {
    __shared__ double sh[4 * 4];
    const BlockRef ref = block_ref(b.blk_cell, b.pfx_cell, b.nd, blockIdx.x);
    if (ref.d < 0) return;
    const int d = ref.d;
    const DesignDev& D = b.d[d];
    const DevState* st = D.st;
    if (st->done) return;
    if (MODE == 1 && !st->need_fix) return;
    const int blk = ref.blk;
    const int i = blk * kBlock + threadIdx.x;
    const int it = st->it, par = it & 1;
    double part[4] = {0, 0, 0, 0};
    if (i < D.n_mov) {
        // the cell's own operands of the combine step are loaded first, beside
        // its position, so their latency overlaps the field loads (dpos and the
        // previous gradient hold stale values at it = 0 and are then unused)
        double2 wl = make_double2(0.0, 0.0);
        double pw = 0.0, dpv[2] = {0.0, 0.0}, gpv[2] = {0.0, 0.0};
        if (MODE == 0) {
            const typename Vec2<T>::type w = ((const typename Vec2<T>::type*)D.gwl)[i];
            wl = make_double2((double)w.x, (double)w.y);
            pw = (double)((const T*)D.pinw)[i];
#pragma unroll
            for (int ax = 0; ax < 2; ax++) {
                dpv[ax] = (double)((const T*)D.dpos[ax])[i];
                gpv[ax] = (double)((const T*)D.g[par ^ 1][ax])[i];
            }
        }
        double gden[2] = {0, 0};
        if (MODE != 1) {
            const typename Vec2<T>::type l = ((const typename Vec2<T>::type*)D.look)[i];
            const double cx = (double)l.x, cy = (double)l.y;
            const BinBox r = bin_box(D, cx, cy, (double)D.hs[i]);
            typedef typename Vec2<T>::type V;
            const V* E = (const V*)D.field;
            double fx = 0, fy = 0, wt = 0;
            if (r.bx1 - r.bx0 <= 3 && r.by1 - r.by0 <= 3) {
                // footprints of up to 3x3 bins (all of them at areas <= 4 and
                // bins no smaller than tiles): the nine field pairs are loaded
                // together, then added in the same row-major order as below
                T e0[9], e1[9];
                if constexpr (std::is_same<T, float>::value) {
                    // FAST path: a row's three pairs in two loads, the first
                    // two pairs as one 16-byte load (at most 6 addresses per
                    // cell instead of 9: the gather is bound by the scattered
                    // addresses the L1 takes per clock).  Every address is
                    // inside the field: rows past the plane are clamped to
                    // its last row and the pair to the row's last two bins
                    // (bx0 on the last bin is then the pair's upper half);
                    // bins outside the footprint are zeroed below.
                    typedef float f4a8 __attribute__((ext_vector_type(4), aligned(8)));
                    const int i0 = min(r.bx0, D.Bx - 2), i2 = min(r.bx0 + 2, D.Bx - 1);
                    const bool hi = i0 != r.bx0;
#pragma unroll
                    for (int j = 0; j < 3; j++) {
                        const size_t row = (size_t)min(r.by0 + j, D.By - 1) * D.Bx;
                        // the third row and column only where the footprint reaches them
                        f4a8 pr = {0.f, 0.f, 0.f, 0.f};
                        V t = {0.f, 0.f};
                        if (j < 2 || r.by0 + 2 < r.by1) {
                            pr = *(const f4a8*)&E[row + i0];
                            if (r.bx0 + 2 < r.bx1) t = E[row + i2];
                        }
                        e0[3 * j] = hi ? pr.z : pr.x;
                        e1[3 * j] = hi ? pr.w : pr.y;
                        e0[3 * j + 1] = pr.z;
                        e1[3 * j + 1] = pr.w;
                        e0[3 * j + 2] = t.x;
                        e1[3 * j + 2] = t.y;
                    }
#pragma unroll
                    for (int q = 0; q < 9; q++) {
                        const bool in = r.by0 + q / 3 < r.by1 && r.bx0 + q % 3 < r.bx1;
                        e0[q] = in ? e0[q] : T(0);
                        e1[q] = in ? e1[q] : T(0);
                    }
                } else {
#pragma unroll
                    for (int q = 0; q < 9; q++) {
                        const int by = r.by0 + q / 3, bx = r.bx0 + q % 3;
                        const bool in = by < r.by1 && bx < r.bx1;
                        const size_t idx = in ? (size_t)by * D.Bx + bx : 0;
                        const V e = E[idx];
                        e0[q] = in ? e.x : T(0);
                        e1[q] = in ? e.y : T(0);
                    }
                }
                double oxs[3], oys[3];  // w = ov_area(r, bx, by), formed per axis
#pragma unroll
                for (int u = 0; u < 3; u++) {
                    oxs[u] = ov_x(r, r.bx0 + u);
                    oys[u] = ov_y(r, r.by0 + u);
                }
#pragma unroll
                for (int q = 0; q < 9; q++) {
                    const int by = r.by0 + q / 3, bx = r.bx0 + q % 3;
                    if (by < r.by1 && bx < r.bx1) {
                        double w = oxs[q % 3] * oys[q / 3];
                        if (w > 0.0) {
                            fx += w * (double)e0[q];
                            fy += w * (double)e1[q];
                            wt += w;
                        }
                    }
                }
            } else {
                for (int by = r.by0; by < r.by1; by++)
                    for (int bx = r.bx0; bx < r.bx1; bx++) {
                        double w = ov_area(r, bx, by);
                        if (w > 0.0) {
                            const V e = E[(size_t)by * D.Bx + bx];
                            fx += w * (double)e.x;
                            fy += w * (double)e.y;
                            wt += w;
                        }
                    }
            }
            if (wt > 1e-10) {
                if constexpr (std::is_same<T, float>::value) {
                    // FAST: one hardware reciprocal (1 ulp of FP32) for both axes
                    const double iw = (double)__builtin_amdgcn_rcpf((float)wt);
                    gden[0] = fx * iw;
                    gden[1] = fy * iw;
                } else {
                    gden[0] = fx / wt;
                    gden[1] = fy / wt;
                }
            }
        } else {
            gden[0] = (double)((const T*)D.g[par][0])[i];
            gden[1] = (double)((const T*)D.g[par][1])[i];
        }
        if (MODE == 2) {
            ex_out[i] = gden[0];
            ey_out[i] = gden[1];
        } else if (MODE == 0 && !st->lam_init) {
            // penalty not yet initialised: park the density gradient, norms only
            const double w0 = wl.x, w1 = wl.y;
            ((T*)D.g[par][0])[i] = (T)gden[0];
            ((T*)D.g[par][1])[i] = (T)gden[1];
            part[0] = w0 * w0 + w1 * w1;
            part[1] = gden[0] * gden[0] + gden[1] * gden[1];
        } else {
            const double lam = MODE == 0 ? sched(D.cfg, st->lam) : st->lam;
            if (MODE != 0) {  // the fix pass reads its operands here
                const typename Vec2<T>::type w = ((const typename Vec2<T>::type*)D.gwl)[i];
                wl = make_double2((double)w.x, (double)w.y);
                pw = (double)((const T*)D.pinw)[i];
            }
            double pc = pw + lam;  // mod.rs:152-156
            pc = pc < 1.0 ? 1.0 : pc;
            const double gwa[2] = {wl.x, wl.y};
            // FAST: the preconditioner's reciprocal (pc >= 1) once, in FP32 (1 ulp)
            const double ipc = std::is_same<T, float>::value ? (double)__builtin_amdgcn_rcpf((float)pc) : 0.0;
#pragma unroll
            for (int ax = 0; ax < 2; ax++) {
                const double gw = gwa[ax];
                T g = std::is_same<T, float>::value ? (T)((gw + lam * gden[ax]) * ipc)
                                                    : (T)((gw + lam * gden[ax]) / pc);  // mod.rs:146-149
                ((T*)D.g[par][ax])[i] = g;
                if (it > 0) {
                    double dx = MODE == 0 ? dpv[ax] : (double)((const T*)D.dpos[ax])[i];
                    double dg = (double)g - (MODE == 0 ? gpv[ax] : (double)((const T*)D.g[par ^ 1][ax])[i]);
                    part[2 * ax + 0] = dx * dx;
                    part[2 * ax + 1] = dx * dg;
                }
            }
        }
    }
    if (MODE == 2) return;
    block_sum<4>(part, sh);
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < 4; q++) D.part_cell[(size_t)blk * kCellParts + q] = part[q];
}

// ------------------------------------------------------------ control ---

__device__ inline void bsum_range(const double* p, int n, int stride, int off, double& out,
                                  double* sh) {
    // fixed-order strided sum then a fixed tree: deterministic
    double v[1] = {0.0};
    for (int k = threadIdx.x; k < n; k += blockDim.x) v[0] += p[(size_t)k * stride + off];
    block_sum<1>(v, sh);
    __syncthreads();
    if (threadIdx.x == 0) sh[16] = v[0];
    __syncthreads();
    out = sh[16];
    __syncthreads();
}

// NV columns of a strided partials array in one pass, each summed in the
// same order as bsum_range (so bit for bit the same); the sums are valid in
// thread 0 only.
template <int NV>
__device__ inline void bsum_cols(const double* p, int n, int stride, int off0, double (&out)[NV],
                                 double* sh) {
#pragma unroll
    for (int q = 0; q < NV; q++) out[q] = 0.0;
    for (int k = threadIdx.x; k < n; k += blockDim.x)
#pragma unroll
        for (int q = 0; q < NV; q++) out[q] += p[(size_t)k * stride + off0 + q];
    block_sum<NV>(out, sh);
    __syncthreads();
}

// Reduce the partials of the previous update (step norms, HPWL) into its log record.
__device__ inline void reduce_prev(const DesignDev& D, DevState* st, double* sh) {
    if (!st->prev_pending) return;
    double sq[2];
    bsum_cols<2>(D.part_cell, D.n_cell_blocks, kCellParts, 4, sq, sh);
    double h = 0.0;
    bool hp = st->prev_check && D.cfg.log_hpwl;
    if (hp) bsum_range(D.part_net, D.n_wa_waves + D.n_hf, 1, 0, h, sh);
    if (threadIdx.x == 0) {
        const int k = st->it - 1;
        if (k >= 0 && k < D.cfg.max_iters) {
            D.log[k].step_norm_x = sqrt(sq[0]);
            D.log[k].step_norm_y = sqrt(sq[1]);
            D.log[k].hpwl = hp ? h : 0.0;
        }
        if (hp) st->hpwl = h;
        st->prev_pending = 0;
    }
    __syncthreads();
}

// After the gather (mod.rs:131-143, 159-166): overflow, penalty schedule or
// initialisation, BB step sizes, the iteration's log record.
__device__ inline void ctrl_after_gather(const DesignDev& D, DevState* st, double* sh) {
    const DevCfg& c = D.cfg;
    const int k = st->it;
    reduce_prev(D, st, sh);
    double bins[2], cells[4];
    bsum_cols<2>(D.part_bin, D.n_ov_parts, kBinParts, 0, bins, sh);
    bsum_cols<4>(D.part_cell, D.n_cell_blocks, kCellParts, 0, cells, sh);
    if (threadIdx.x != 0) return;
    const double tot = bins[0], ovs = bins[1];
    const double p0 = cells[0], p1 = cells[1], p2 = cells[2], p3 = cells[3];
    const double ov = tot < 1e-30 ? 0.0 : ovs / tot;  // density.rs:79-86
    st->ov = ov;
    double wln = 0.0, dnn = 0.0;
    if (st->lam_init) {
        st->lam = sched(c, st->lam);
        if (k > 0) {  // mod.rs:159-166 with nesterov.rs:199-218
            if (fabs(p1) >= 1e-30) st->s[0] = fmin(fmax(fabs(p0 / p1), 1e-4), 1.0);
            if (fabs(p3) >= 1e-30) st->s[1] = fmin(fmax(fabs(p2 / p3), 1e-4), 1.0);
        }
        st->need_fix = 0;
    } else {
        wln = sqrt(p0);  // common.rs:418-425
        dnn = sqrt(p1);
        if (dnn > 1e-30) {  // mod.rs:132-138
            st->lam = c.eta * wln / dnn;
            st->lam_init = 1;
        }
        st->need_fix = 1;
    }
    st->wl_norm = wln;
    st->den_norm = dnn;
    st->check = c.stop_every_iter || (k % c.interval == 0) || (k == c.max_iters - 1);
    if (c.check_nonfinite && !(isfinite(ov) && isfinite(st->lam))) st->nonfinite = 1;
    ep_iter_rec& r = D.log[k];
    r.iter = k;
    r.checked = st->check;
    r.overflow = ov;
    r.lambda = st->lam;
    r.wl_norm = wln;
    r.den_norm = dnn;
    r.step_norm_x = r.step_norm_y = 0.0;
    r.hpwl = 0.0;
}

// End of the control step (nesterov.rs:41-47, mod.rs:176-195): BB steps of
// an iteration whose penalty was just initialised, FISTA sequence, stop rule.
__device__ inline void ctrl_end(const DesignDev& D, DevState* st, double* sh) {
    const DevCfg& c = D.cfg;
    const int k = st->it;
    if (st->need_fix) {
        double cells[4];
        bsum_cols<4>(D.part_cell, D.n_cell_blocks, kCellParts, 0, cells, sh);
        if (threadIdx.x == 0 && k > 0) {
            if (fabs(cells[1]) >= 1e-30) st->s[0] = fmin(fmax(fabs(cells[0] / cells[1]), 1e-4), 1.0);
            if (fabs(cells[3]) >= 1e-30) st->s[1] = fmin(fmax(fabs(cells[2] / cells[3]), 1e-4), 1.0);
        }
    }
    if (threadIdx.x != 0) return;
    st->need_fix = 0;
    // FISTA sequence (nesterov.rs:41-47 and :75-82): the step of iteration k
    // uses m_k = (a_k - 1) / a_{k+1}; the next look-ahead uses m_{k+1}.
    const double a = st->a;
    const double an = (1.0 + sqrt(1.0 + 4.0 * a * a)) / 2.0;
    const double ann = (1.0 + sqrt(1.0 + 4.0 * an * an)) / 2.0;
    st->m = (a - 1.0) / an;
    st->m_next = (an - 1.0) / ann;
    st->a = an;
    ep_iter_rec& r = D.log[k];
    r.step_size_x = st->s[0];
    r.step_size_y = st->s[1];
    bool stop = false;
    if (st->check && k >= c.min_iters && st->ov < c.overflow_target) {  // mod.rs:192-195
        st->converged = 1;
        stop = true;
    }
    if (k + 1 >= c.max_iters) stop = true;
    if (st->nonfinite) stop = true;
    st->run_update = 1;
    st->prev_pending = 1;
    st->prev_check = st->check;
    st->it = k + 1;
    st->n_log = k + 1;
    if (stop) {
        st->done = 1;
        st->t_done = wall_clock64();
    }
}

// which = 0: after the gather; 1: after the fix (end of control); 4: both
// in one launch (iterations without a fix pass); 2: finalize pending log
// partials; 3: reduce a final HPWL pass.  One block per design.
__global__ __launch_bounds__(256) void k_ctrl(BatchDev b, int which) {
    __shared__ double sh[32];
    const DesignDev& D = b.d[blockIdx.x];
    DevState* st = D.st;
    if (which == 2) {
        reduce_prev(D, st, sh);
        return;
    }
    if (which == 3) {
        double h;
        bsum_range(D.part_net, D.n_wa_waves + D.n_hf, 1, 0, h, sh);
        if (threadIdx.x == 0) st->final_hpwl = h;
        return;
    }
    if (st->done) {
        if (which != 0 && threadIdx.x == 0) st->run_update = 0;
        return;
    }
    if (which == 0 || which == 4) ctrl_after_gather(D, st, sh);
    if (which == 4) __syncthreads();  // thread 0's state writes, seen by the block
    if (which == 1 || which == 4) ctrl_end(D, st, sh);
}

// ------------------------------------------------------------ update ----

template <typename T, bool FIX>
__global__ __launch_bounds__(256) void k_update(BatchDev b) {
    __shared__ double sh[4 * 2];
    const BlockRef ref = block_ref(b.blk_cell, b.pfx_cell, b.nd, blockIdx.x);
    if (ref.d < 0) return;
    const int d = ref.d;
    const DesignDev& D = b.d[d];
    const DevState* st = D.st;
    if (!st->run_update) return;
    const int blk = ref.blk;
    const int i = blk * kBlock + threadIdx.x;
    const int par = (st->it - 1) & 1;
    double sq[2] = {0, 0};
    if (i < D.n_mov) {
        const T m = (T)st->m, mn = (T)st->m_next;
        // both axes' operands are loaded before either axis stores (the
        // stores could alias the other axis's loads as far as the compiler
        // knows, which would order each axis's loads after the other's stores)
        T xa[2], dxa[2], ga[2];
#pragma unroll
        for (int ax = 0; ax < 2; ax++) {
            xa[ax] = ((const T*)D.pos[ax])[i];
            dxa[ax] = ((const T*)D.dpos[ax])[i];
            ga[ax] = ((const T*)D.g[par][ax])[i];
        }
        T lk[2];
#pragma unroll
        for (int ax = 0; ax < 2; ax++) {
            const T hi = (T)(ax == 0 ? D.max_x : D.max_y);
            const T s = (T)st->s[ax];
            T* P = (T*)D.pos[ax];
            T* DP = (T*)D.dpos[ax];
            const T x = xa[ax], dx = dxa[ax];
            const T g = ga[ax];
            const T yv = x + m * dx;  // nesterov.rs:55: y = x + m (x - x_prev)
            const T xn = yv - s * g;  // nesterov.rs:56
            const double dl = (double)xn - (double)x;
            sq[ax] = dl * dl;
            const T xnc = xn < T(0) ? T(0) : (xn > hi ? hi : xn);  // nesterov.rs:67-72
            const T xpc = x < T(0) ? T(0) : (x > hi ? hi : x);
            P[i] = xnc;
            const T ndx = xnc - xpc;
            DP[i] = ndx;
            T l = xnc + mn * ndx;  // nesterov.rs:95-101 then common.rs:410-415
            l = l < T(0) ? T(0) : (l > hi ? hi : l);
            lk[ax] = l;
        }
        typename Vec2<T>::type v;
        v.x = lk[0];
        v.y = lk[1];
        ((typename Vec2<T>::type*)D.look)[i] = v;
    }
    block_sum<2>(sq, sh);
    if (threadIdx.x == 0) {
        D.part_cell[(size_t)blk * kCellParts + 4] = sq[0];
        D.part_cell[(size_t)blk * kCellParts + 5] = sq[1];
    }
}

// Reset a design to its initial positions (NesterovSolver::set_positions,
// nesterov.rs:114-118; first look-ahead has zero momentum, nesterov.rs:75-82).
template <typename T, bool FIX>
__global__ __launch_bounds__(256) void k_reset(BatchDev b) {
    const BlockRef ref = block_ref(b.blk_cell, b.pfx_cell, b.nd, blockIdx.x);
    if (ref.d < 0) return;
    const int d = ref.d;
    const DesignDev& D = b.d[d];
    const int i = ref.blk * kBlock + threadIdx.x;
    if (i >= D.n_mov) return;
    if (i == 0) D.st->t_start = wall_clock64();  // the placement's start, for its latency
    const uint32_t c0 = D.perm0[i];
    D.perm[0][i] = c0;
    D.sorted_hs[0][i] = D.hs[c0];
    T lk[2];
#pragma unroll
    for (int ax = 0; ax < 2; ax++) {
        const T hi = (T)(ax == 0 ? D.max_x : D.max_y);
        const T x = ((const T*)D.x0[ax])[i];
        ((T*)D.pos[ax])[i] = x;
        ((T*)D.dpos[ax])[i] = T(0);
        lk[ax] = x < T(0) ? T(0) : (x > hi ? hi : x);
    }
    typename Vec2<T>::type v;
    v.x = lk[0];
    v.y = lk[1];
    ((typename Vec2<T>::type*)D.look)[i] = v;
}

// Evaluation helper: the WA gradient per cell (the gather's slot reduction).
template <typename T>
__global__ __launch_bounds__(256) void k_wa_reduce(BatchDev b, double* gx, double* gy) {
    const BlockRef ref = block_ref(b.blk_cell, b.pfx_cell, b.nd, blockIdx.x);
    if (ref.d < 0) return;
    const int d = ref.d;
    const DesignDev& D = b.d[d];
    const int i = ref.blk * kBlock + threadIdx.x;
    if (i >= D.n_mov) return;
    double a, c;
    gwl_of<T>(D, i, a, c);
    gx[i] = a;
    gy[i] = c;
}

// Evaluation helper: integer bin ranges per cell.
__global__ __launch_bounds__(256) void k_bin_ranges(BatchDev b, int32_t* out) {
    const BlockRef ref = block_ref(b.blk_cell, b.pfx_cell, b.nd, blockIdx.x);
    if (ref.d < 0) return;
    const int d = ref.d;
    const DesignDev& D = b.d[d];
    const int i = ref.blk * kBlock + threadIdx.x;
    if (i >= D.n_mov) return;
    double cx, cy;
    if (D.prec == EP_PREC_FP64) {
        cx = ((const double*)D.look)[2 * (size_t)i];
        cy = ((const double*)D.look)[2 * (size_t)i + 1];
    } else {
        cx = (double)((const float*)D.look)[2 * (size_t)i];
        cy = (double)((const float*)D.look)[2 * (size_t)i + 1];
    }
    BinBox r = bin_box(D, cx, cy, (double)D.hs[i]);
    out[4 * i + 0] = r.bx0;
    out[4 * i + 1] = r.bx1;
    out[4 * i + 2] = r.by0;
    out[4 * i + 3] = r.by1;
}

// ----------------------------------------------------------- launchers --

void launch_wa(const BatchDev& b, int prec, int wa_blocks, int hf_blocks, hipStream_t s) {
    if (wa_blocks > 0) {
        if (prec == EP_PREC_FP64)
            hipLaunchKernelGGL((k_wa<double, false, false>), dim3(wa_blocks), dim3(256), 0, s, b, 0);
        else
            hipLaunchKernelGGL((k_wa<float, true, false>), dim3(wa_blocks), dim3(256), 0, s, b, 0);
    }
    if (hf_blocks > 0) {
        if (prec == EP_PREC_FP64)
            hipLaunchKernelGGL((k_wa_hf<double, false, false>), dim3(hf_blocks), dim3(kHfThreads), 0, s, b, 0);
        else
            hipLaunchKernelGGL((k_wa_hf<float, true, false>), dim3(hf_blocks), dim3(kHfThreads), 0, s, b, 0);
    }
}

void launch_hpwl(const BatchDev& b, int prec, int wa_blocks, int hf_blocks, int final_pass,
                 hipStream_t s) {
    if (wa_blocks > 0) {
        if (prec == EP_PREC_FP64)
            hipLaunchKernelGGL((k_wa<double, false, true>), dim3(wa_blocks), dim3(256), 0, s, b, final_pass);
        else
            hipLaunchKernelGGL((k_wa<float, true, true>), dim3(wa_blocks), dim3(256), 0, s, b, final_pass);
    }
    if (hf_blocks > 0) {
        if (prec == EP_PREC_FP64)
            hipLaunchKernelGGL((k_wa_hf<double, false, true>), dim3(hf_blocks), dim3(kHfThreads), 0, s, b, final_pass);
        else
            hipLaunchKernelGGL((k_wa_hf<float, true, true>), dim3(hf_blocks), dim3(kHfThreads), 0, s, b, final_pass);
    }
}

void launch_scatter(const BatchDev& b, int prec, int scat_blocks, hipStream_t s) {
    if (scat_blocks <= 0) return;
    if (prec == EP_PREC_FP64)
        hipLaunchKernelGGL((k_scatter<double, false>), dim3(scat_blocks), dim3(256), 0, s, b);
    else
        hipLaunchKernelGGL((k_scatter<float, true>), dim3(scat_blocks), dim3(256), 0, s, b);
}

void launch_wa_reduce(const BatchDev& b, int prec, int cell_blocks, double* gx, double* gy, hipStream_t s) {
    if (cell_blocks <= 0) return;
    if (prec == EP_PREC_FP64)
        hipLaunchKernelGGL(k_wa_reduce<double>, dim3(cell_blocks), dim3(256), 0, s, b, gx, gy);
    else
        hipLaunchKernelGGL(k_wa_reduce<float>, dim3(cell_blocks), dim3(256), 0, s, b, gx, gy);
}

void launch_dens_final(const BatchDev& b, int prec, int bin_blocks, hipStream_t s) {
    if (bin_blocks <= 0) return;
    if (prec == EP_PREC_FP64)
        hipLaunchKernelGGL((k_dens_final<double, false>), dim3(bin_blocks), dim3(256), 0, s, b);
    else
        hipLaunchKernelGGL((k_dens_final<float, true>), dim3(bin_blocks), dim3(256), 0, s, b);
}

void launch_wl_cell(const BatchDev& b, int prec, int cell_blocks, hipStream_t s) {
    if (cell_blocks <= 0) return;
    if (prec == EP_PREC_FP64)
        hipLaunchKernelGGL(k_wl_cell<double>, dim3(cell_blocks), dim3(256), 0, s, b);
    else
        hipLaunchKernelGGL(k_wl_cell<float>, dim3(cell_blocks), dim3(256), 0, s, b);
}

void launch_gather(const BatchDev& b, int prec, int cell_blocks, int mode, hipStream_t s) {
    if (cell_blocks <= 0) return;
    double* n = nullptr;
    if (prec == EP_PREC_FP64) {
        if (mode == 0) hipLaunchKernelGGL((k_gather<double, false, 0>), dim3(cell_blocks), dim3(256), 0, s, b, n, n);
        else hipLaunchKernelGGL((k_gather<double, false, 1>), dim3(cell_blocks), dim3(256), 0, s, b, n, n);
    } else {
        if (mode == 0) hipLaunchKernelGGL((k_gather<float, true, 0>), dim3(cell_blocks), dim3(256), 0, s, b, n, n);
        else hipLaunchKernelGGL((k_gather<float, true, 1>), dim3(cell_blocks), dim3(256), 0, s, b, n, n);
    }
}

void launch_gather_export(const BatchDev& b, int prec, int cell_blocks, double* gx, double* gy,
                          hipStream_t s) {
    if (cell_blocks <= 0) return;
    if (prec == EP_PREC_FP64)
        hipLaunchKernelGGL((k_gather<double, false, 2>), dim3(cell_blocks), dim3(256), 0, s, b, gx, gy);
    else
        hipLaunchKernelGGL((k_gather<float, true, 2>), dim3(cell_blocks), dim3(256), 0, s, b, gx, gy);
}

void launch_ctrl(const BatchDev& b, int which, hipStream_t s) {
    if (b.nd <= 0) return;
    hipLaunchKernelGGL(k_ctrl, dim3(b.nd), dim3(256), 0, s, b, which);
}

void launch_update(const BatchDev& b, int prec, int cell_blocks, hipStream_t s) {
    if (cell_blocks <= 0) return;
    if (prec == EP_PREC_FP64)
        hipLaunchKernelGGL((k_update<double, false>), dim3(cell_blocks), dim3(256), 0, s, b);
    else
        hipLaunchKernelGGL((k_update<float, true>), dim3(cell_blocks), dim3(256), 0, s, b);
}

void launch_reset(const BatchDev& b, int prec, int cell_blocks, hipStream_t s) {
    if (cell_blocks <= 0) return;
    if (prec == EP_PREC_FP64)
        hipLaunchKernelGGL((k_reset<double, false>), dim3(cell_blocks), dim3(256), 0, s, b);
    else
        hipLaunchKernelGGL((k_reset<float, true>), dim3(cell_blocks), dim3(256), 0, s, b);
}

void launch_tile_sort(const BatchDev& b, int prec, int cell_blocks, hipStream_t s) {
    if (cell_blocks <= 0) return;
    if (prec == EP_PREC_FP64) {
        hipLaunchKernelGGL(k_tile_count<double>, dim3(cell_blocks), dim3(256), 0, s, b);
        hipLaunchKernelGGL(k_tile_scan, dim3(b.nd), dim3(256), 0, s, b);
        hipLaunchKernelGGL(k_tile_fill<double>, dim3(cell_blocks), dim3(256), 0, s, b);
    } else {
        hipLaunchKernelGGL(k_tile_count<float>, dim3(cell_blocks), dim3(256), 0, s, b);
        hipLaunchKernelGGL(k_tile_scan, dim3(b.nd), dim3(256), 0, s, b);
        hipLaunchKernelGGL(k_tile_fill<float>, dim3(cell_blocks), dim3(256), 0, s, b);
    }
}

void launch_scatter_pull(const BatchDev& b, int prec, int tile_blocks, hipStream_t s) {
    if (tile_blocks <= 0) return;
    if (prec == EP_PREC_FP64)
        hipLaunchKernelGGL((k_scatter_pull<double, false>), dim3(tile_blocks), dim3(256), 0, s, b);
    else
        hipLaunchKernelGGL((k_scatter_pull<float, true>), dim3(tile_blocks), dim3(256), 0, s, b);
}

void launch_bin_ranges(const BatchDev& b, int cell_blocks, int32_t* out, hipStream_t s) {
    if (cell_blocks <= 0) return;
    hipLaunchKernelGGL(k_bin_ranges, dim3(cell_blocks), dim3(256), 0, s, b, out);
}

}  // namespace ep
