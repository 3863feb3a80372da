// Spectral Poisson solve as dense transform products (density.rs:96-166).
//
// The reference's three 2-D transforms (DCT-II forward, then DST-III/DCT-III
// and DCT-III/DST-III inverses, each rows then columns) are six dense
// products with constant cosine/sine matrices.  Every product here is
// C = A * Bt^T with A (M x K) and Bt (N x K) both row-major, so both MFMA
// operands are K-contiguous in LDS; a pass whose output feeds the next
// pass's right operand stores it transposed.
//
// FAST engine: BF16 MFMA (v_mfma_f32_32x32x16_bf16) with split-precision
// operands x = hi + lo, three products hi*hi + hi*lo + lo*hi accumulated in
// FP32 (the lo*lo term is below FP32 rounding).  Data operands are split on
// the fly while staging into LDS; the constant matrices are pre-split.
// FP64 engine: an LDS-tiled FMA kernel (also the fallback for shapes that
// are not multiples of the MFMA tile).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>

#include "ep_internal.h"

namespace ep {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// Operand and output pointers come from job descriptors in memory; as
// generic pointers they would compile to flat accesses, which count on the
// LDS/scalar counter too, so every wait for an LDS read or a descriptor field
// would also drain the operand loads in flight.  They are cast to the global
// address space instead.
#define EP_GLOBAL __attribute__((address_space(1)))

template <typename T>
__device__ inline const EP_GLOBAL T* gl(const void* p) {
    return (const EP_GLOBAL T*)p;
}

template <typename T>
__device__ inline EP_GLOBAL T* gw(void* p) {
    return (EP_GLOBAL T*)p;
}


// The spectral scaling and Poisson solve of density.rs:113-153 for
// coefficient v at (ky = row, kx = col), given the spectral factors
// (w, 2(1 - cos w)) of its column (wx, lx) and row (wy, ly).
__device__ inline void poisson(const DesignDev& D, int row, int col, double v, double wx, double lx,
                               double wy, double ly, double& ex, double& ey) {
    double s = 4.0 / ((double)D.Bx * (double)D.By);
    if (col == 0) s *= 0.5;
    if (row == 0) s *= 0.5;
    const double a = v * s;
    const double den = lx + ly;
    ex = 0.0;
    ey = 0.0;
    if (!(den < 1e-30)) {
        const double phi = a / den;
        ex = phi * wx;
        ey = phi * wy;
    }
}

// Store one output element; kEpiPoisson applies poisson() to coefficient
// (ky = row, kx = col).  po: element offset of the output plane (K slice)
// being written.
template <typename T>
__device__ inline void epi_store(const GemmJob& J, const DesignDev& D, int row, int col, double v,
                                 size_t po = 0) {
    if (row >= J.M || col >= J.N) return;
    if (J.epi == kEpiStore) {
        gw<T>(J.C)[po + (size_t)row * J.N + col] = (T)v;
    } else if (J.epi == kEpiStoreT) {
        gw<T>(J.C)[po + (size_t)col * J.M + row] = (T)v;
    } else if (J.epi == kEpiStore2) {
        gw<T>(J.C)[po + 2 * ((size_t)row * J.N + col)] = (T)v;
    } else {
        const EP_GLOBAL T* sx = gl<T>(D.spec_x);  // [kx] -> (w, 2(1 - cos w))
        const EP_GLOBAL T* sy = gl<T>(D.spec_y);
        double ex, ey;
        poisson(D, row, col, v, (double)sx[2 * col], (double)sx[2 * col + 1], (double)sy[2 * row],
                (double)sy[2 * row + 1], ex, ey);
        gw<T>(J.C)[po + (size_t)row * J.N + col] = (T)ex;
        gw<T>(J.C2)[po + (size_t)row * J.N + col] = (T)ey;
    }
}

// --------------------------------------------------------------- FMA -----

template <typename T>
__global__ __launch_bounds__(256) void k_gemm_fma(const GemmJob* jobs, const int32_t* bjob, const DesignDev* dd) {
    __shared__ T sA[16][64 + 1];
    __shared__ T sB[16][64 + 1];
    const int j = gl<int32_t>(bjob)[2 * blockIdx.x];
    if (j < 0) return;
    const int tb = gl<int32_t>(bjob)[2 * blockIdx.x + 1];
    const GemmJob& J = jobs[j];
    if (J.st->done) return;
    const DesignDev& D = dd[J.design];
    const int t = tb - J.tile_start;
    const int m0 = (t / J.tiles_n) * 64, n0 = (t % J.tiles_n) * 64;
    const int tid = threadIdx.x, ty = tid >> 4, tx = tid & 15;
    const EP_GLOBAL T* A = gl<T>(J.A);
    const EP_GLOBAL T* B = gl<T>(J.B);
    T acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int q = 0; q < 4; q++) acc[i][q] = T(0);
    for (int k0 = 0; k0 < J.K; k0 += 16) {
        for (int e = tid; e < 64 * 16; e += 256) {
            const int r = e >> 4, kk = e & 15;
            const int k = k0 + kk;
            sA[kk][r] = (m0 + r < J.M && k < J.K) ? A[(size_t)(m0 + r) * J.K + k] : T(0);
            sB[kk][r] = (n0 + r < J.N && k < J.K) ? B[(size_t)(n0 + r) * J.K + k] : T(0);
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; kk++) {
            T a[4], bb[4];
#pragma unroll
            for (int i = 0; i < 4; i++) a[i] = sA[kk][ty + 16 * i];
#pragma unroll
            for (int q = 0; q < 4; q++) bb[q] = sB[kk][tx + 16 * q];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int q = 0; q < 4; q++) acc[i][q] += a[i] * bb[q];
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int q = 0; q < 4; q++) epi_store<T>(J, D, m0 + ty + 16 * i, n0 + tx + 16 * q, (double)acc[i][q]);
}

// -------------------------------------------------------------- MFMA -----

__device__ inline void split8(const float* v, bf16x8& hi, bf16x8& lo) {
#pragma unroll
    for (int q = 0; q < 8; q++) {
        __bf16 h = (__bf16)v[q];
        hi[q] = h;
        lo[q] = (__bf16)(v[q] - (float)h);
    }
}

// One thread's share of one K stage (64 rows x TK) of each operand: TK / 32
// chunks of 8 consecutive K values of row tid / 4 (chunk c at K offset
// 8 (tid % 4) + 32 c).  The data operand stays FP32 in registers until the
// stage is written to LDS, where it is split into BF16 hi/lo: a wait for the
// load then sits where the stage is consumed, not where it is issued.  The
// constant operand arrives pre-split.
template <int TK>
struct Stage {
    f32x4 d0[TK / 32], d1[TK / 32];   // data: K values 0-3 and 4-7 of each chunk
    bf16x8 hi[TK / 32], lo[TK / 32];  // constant
};

// A data operand written by a split pass is the sum of its `parts` planes
// (pstride elements apart), added in plane order (SUM only).
template <int TK, bool SUM>
__device__ inline void load_stage(Stage<TK>& s, const float* pd, int rd, const __bf16* ch,
                                  const __bf16* cl, int rc, int k0, int K, int parts,
                                  int64_t pstride) {
    const int row = threadIdx.x >> 2, kc = (threadIdx.x & 3) * 8;
#pragma unroll
    for (int c = 0; c < TK / 32; c++) {
        const size_t od = (size_t)(rd + row) * K + k0 + kc + 32 * c;
        const EP_GLOBAL f32x4* src = gl<f32x4>(pd + od);
        s.d0[c] = src[0];
        s.d1[c] = src[1];
        if (SUM)
            for (int q = 1; q < parts; q++) {
                const EP_GLOBAL f32x4* sq = gl<f32x4>(pd + od + q * pstride);
                s.d0[c] += sq[0];
                s.d1[c] += sq[1];
            }
        const size_t oc = (size_t)(rc + row) * K + k0 + kc + 32 * c;
        s.hi[c] = *gl<bf16x8>(ch + oc);
        s.lo[c] = *gl<bf16x8>(cl + oc);
    }
}

// LDS row layout of a stage.  TK = 32: rows of 64 bytes (four 16-byte
// chunks of 8 BF16) with chunk c of row r stored at chunk c ^ ((r / 4) % 4),
// so that in units of 16 bytes (sixteen per pass over the 64 banks) the
// stage writes (four lanes per row, four consecutive rows per 16 lanes) and
// the MFMA operand reads (16 consecutive rows, one chunk) each touch 16
// distinct units: no bank conflicts.  Otherwise rows are padded by 8 BF16.
template <int TK>
constexpr int lds_row() { return TK == 32 ? TK : TK + 8; }

template <int TK>
__device__ inline int lds_at(int row, int k) {  // element offset of (row, k) within its row
    if constexpr (TK == 32) return (((k >> 3) ^ ((row >> 2) & 3)) << 3) | (k & 7);
    else return k;
}

template <int TK>
__device__ inline void put_stage(__bf16 (*sd)[64][lds_row<TK>()], __bf16 (*sc)[64][lds_row<TK>()],
                                 const Stage<TK>& s) {
    const int row = threadIdx.x >> 2, kc = (threadIdx.x & 3) * 8;
#pragma unroll
    for (int c = 0; c < TK / 32; c++) {
        const float v[8] = {s.d0[c][0], s.d0[c][1], s.d0[c][2], s.d0[c][3],
                            s.d1[c][0], s.d1[c][1], s.d1[c][2], s.d1[c][3]};
        bf16x8 hi, lo;
        split8(v, hi, lo);
        const int at = lds_at<TK>(row, kc + 32 * c);
        *(bf16x8*)&sd[0][row][at] = hi;
        *(bf16x8*)&sd[1][row][at] = lo;
        *(bf16x8*)&sc[0][row][at] = s.hi[c];
        *(bf16x8*)&sc[1][row][at] = s.lo[c];
    }
}

// The three split products hi*hi + hi*lo + lo*hi of one LDS stage into the
// wave's 32 x 32 quadrant.
template <int TK>
__device__ inline void mma_stage(f32x16& acc, __bf16 (*sa)[64][lds_row<TK>()],
                                 __bf16 (*sb)[64][lds_row<TK>()], int ra, int rb, int kh) {
#pragma unroll
    for (int kk = 0; kk < TK; kk += 16) {
        const int aa = lds_at<TK>(ra, kk + kh), ab = lds_at<TK>(rb, kk + kh);
        const bf16x8 ah = *(const bf16x8*)&sa[0][ra][aa];
        const bf16x8 al = *(const bf16x8*)&sa[1][ra][aa];
        const bf16x8 bh = *(const bf16x8*)&sb[0][rb][ab];
        const bf16x8 bl = *(const bf16x8*)&sb[1][rb][ab];
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc, 0, 0, 0);
    }
}

// 64 x 64 output tile per block, four waves each owning a 32 x 32 quadrant,
// K in stages of TK through NBUF LDS buffers (rows laid out by lds_at).  Two
// register sets alternate (the loop is unrolled by two, so no register copy
// waits on a load in flight): while a stage is multiplied, the loads of the
// next two stages are outstanding.  Past the last stage the loads repeat the
// last stage (in bounds, unused).  TK = 64 with one buffer halves the chain
// of stages in about the LDS of TK = 32 with two (40 and 32 KB: four
// resident blocks per CU).  ADATA: the data operand is A (F1, I1), else B (F2, I2).
// Requires M, N multiples of 64 and K a multiple of TK (checked on the host).
template <int TK, int NBUF, bool ADATA, bool SUM>
__global__ __launch_bounds__(256) void k_gemm_mfma(const GemmJob* jobs, const int32_t* bjob, const DesignDev* dd) {
    __shared__ __bf16 sA[NBUF][2][64][lds_row<TK>()];  // [buffer][hi/lo][row][k (laid out by lds_at)]
    __shared__ __bf16 sB[NBUF][2][64][lds_row<TK>()];
    constexpr int b1 = NBUF - 1;  // the second buffer (the first again when single)
    const int j = gl<int32_t>(bjob)[2 * blockIdx.x];
    if (j < 0) return;
    const int tb = gl<int32_t>(bjob)[2 * blockIdx.x + 1];
    const GemmJob& J = jobs[j];
    if (J.st->done) return;
    const DesignDev& D = dd[J.design];
    // block -> (K slice, tile): consecutive blocks take different tiles
    const int tiles = ((J.M + 63) / 64) * J.tiles_n;
    const int tt = tb - J.tile_start;
    const int ks = tt / tiles, t = tt - ks * tiles;
    const int kper = J.K / J.ksplit, kbeg = ks * kper;
    const int m0 = (t / J.tiles_n) * 64, n0 = (t % J.tiles_n) * 64;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int ra = (wv >> 1) * 32 + (lane & 31), rb = (wv & 1) * 32 + (lane & 31);
    const int kh = (lane >> 5) * 8;
    const float* pd = (const float*)(ADATA ? J.A : J.B);
    const __bf16* ch = (const __bf16*)(ADATA ? J.B_hi : J.A_hi);
    const __bf16* cl = (const __bf16*)(ADATA ? J.B_lo : J.A_lo);
    const int rd = ADATA ? m0 : n0, rc = ADATA ? n0 : m0;
    const int parts = ADATA ? J.a_parts : J.b_parts;
    const int64_t pstride = ADATA ? J.a_pstride : J.b_pstride;
    __bf16(*sD)[2][64][lds_row<TK>()] = ADATA ? sA : sB;
    __bf16(*sC)[2][64][lds_row<TK>()] = ADATA ? sB : sA;
    f32x16 acc;
#pragma unroll
    for (int q = 0; q < 16; q++) acc[q] = 0.0f;
    const int stages = kper / TK;
    const int klast = kbeg + (stages - 1) * TK;
    auto kat = [&](int st) { return min(kbeg + st * TK, klast); };
    Stage<TK> s0, s1;
    load_stage<TK, SUM>(s0, pd, rd, ch, cl, rc, kbeg, J.K, parts, pstride);
    load_stage<TK, SUM>(s1, pd, rd, ch, cl, rc, kat(1), J.K, parts, pstride);
    put_stage<TK>(sD[0], sC[0], s0);
    __syncthreads();
    // Every path through the loop body consumes the loads it carries in (a
    // load left pending across the back edge would make the wait inserted at
    // the loop head drain every load in flight), and the scheduling barriers
    // keep each half's loads at its top, ahead of the split arithmetic that
    // waits for the previous half's loads.
    // With a single LDS buffer, a barrier separates each stage's products
    // from the put that overwrites it.
    for (int st = 0; st + 1 < stages; st += 2) {
        // buffer 0 holds stage st, s1 carries stage st + 1
        load_stage<TK, SUM>(s0, pd, rd, ch, cl, rc, kat(st + 2), J.K, parts, pstride);
        __builtin_amdgcn_sched_barrier(0);
        mma_stage<TK>(acc, sA[0], sB[0], ra, rb, kh);
        if (NBUF == 1) __syncthreads();
        put_stage<TK>(sD[b1], sC[b1], s1);
        __syncthreads();
        // buffer b1 holds stage st + 1, s0 carries stage st + 2
        load_stage<TK, SUM>(s1, pd, rd, ch, cl, rc, kat(st + 3), J.K, parts, pstride);
        __builtin_amdgcn_sched_barrier(0);
        mma_stage<TK>(acc, sA[b1], sB[b1], ra, rb, kh);
        if (NBUF == 1) __syncthreads();
        put_stage<TK>(sD[0], sC[0], s0);
        __syncthreads();
    }
    if (stages & 1) mma_stage<TK>(acc, sA[0], sB[0], ra, rb, kh);  // buffer 0 holds the last stage
    // accumulator layout: element q of lane l is (row 8*(q/4) + 4*(l/32) + q%4, col l%32)
    const size_t po = (size_t)ks * (size_t)J.c_pstride;
    const int wm = wv >> 1, wn = wv & 1;
    const int col = n0 + wn * 32 + (lane & 31);
    if (J.epi == kEpiPoisson) {
        // the spectral factors of the lane's column and sixteen rows are all
        // loaded before the first store (the stores could alias them as far
        // as the compiler knows, which would order each load after the
        // previous element's store)
        const EP_GLOBAL f32x2* sx = gl<f32x2>(D.spec_x);
        const EP_GLOBAL f32x2* sy = gl<f32x2>(D.spec_y);
        const f32x2 fx = sx[col];
        f32x2 fy[16];
#pragma unroll
        for (int q = 0; q < 16; q++) fy[q] = sy[m0 + wm * 32 + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3)];
        // poisson() in FP32 (the outputs are FP32) with the quotient by the
        // hardware reciprocal (1 ulp) instead of an FP64 division per element
        const float s0 = (float)(4.0 / ((double)D.Bx * (double)D.By)) * (col == 0 ? 0.5f : 1.0f);
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int row = m0 + wm * 32 + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
            const float a = acc[q] * (row == 0 ? 0.5f * s0 : s0);
            const float den = fx[1] + fy[q][1];
            const float phi = den < 1e-30f ? 0.f : a * __builtin_amdgcn_rcpf(den);
            gw<float>(J.C)[po + (size_t)row * J.N + col] = phi * fx[0];
            gw<float>(J.C2)[po + (size_t)row * J.N + col] = phi * fy[q][0];
        }
        return;
    }
    if (J.epi == kEpiStoreT) {
        // transposed output (C[col][row]): the tile goes through LDS (free
        // after the last stage; at [col][row ^ col], so that both the
        // quadrant writes and the column reads below hit 64 distinct banks)
        // and each wave then stores whole 256-byte columns, one per
        // instruction, instead of 64 scattered words
        float* T = (float*)&sA[0][0][0][0];
        __syncthreads();  // an odd last stage reads buffer 0 after the loop
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int r = wm * 32 + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3), c = col - n0;
            T[c * 64 + (r ^ c)] = acc[q];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int c = wv * 16 + q;
            gw<float>(J.C)[po + (size_t)(n0 + c) * J.M + m0 + lane] = T[c * 64 + (lane ^ c)];
        }
        return;
    }
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const int row = m0 + wm * 32 + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
        epi_store<float>(J, D, row, col, (double)acc[q], po);
    }
}

// Plane 0 of a split pass's output becomes the sum of its planes, added in
// plane order (one block row per split design).
__global__ __launch_bounds__(256) void k_plane_sum(const PlaneSum* sums) {
    const PlaneSum S = sums[blockIdx.y];
    if ((int)blockIdx.x >= S.blocks) return;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= S.n) return;
    EP_GLOBAL float* p = gw<float>(S.p);
    float v = p[i];
    for (int q = 1; q < S.planes; q++) v += p[i + q * S.n];
    p[i] = v;
}

void launch_plane_sum(const PlaneSum* sums, int n_sums, int max_blocks, hipStream_t s) {
    if (n_sums <= 0 || max_blocks <= 0) return;
    hipLaunchKernelGGL(k_plane_sum, dim3(max_blocks, n_sums), dim3(256), 0, s, sums);
}

// Split-K variant: the four waves of a block each compute the whole 64 x 64
// tile over one quarter of K, with MFMA operands loaded straight from
// global memory (each lane's 8 consecutive K values of one row are 32 bytes
// of FP32 data or 16 + 16 bytes of pre-split BF16); the partial tiles are
// summed in LDS in wave order, so the result is deterministic.  The K chain
// each wave walks is a quarter as long as the single-tile kernel's.
// Requires M, N multiples of 64 and K a multiple of 64.
__device__ inline void frag_direct(int kind, const void* p, const void* plo, int row, int k, int K,
                                   bf16x8& hi, bf16x8& lo) {
    const size_t off = (size_t)row * K + k;
    if (kind == 0) {
        const EP_GLOBAL f32x4* src = gl<f32x4>((const float*)p + off);
        const f32x4 a = src[0], b = src[1];
        const float v[8] = {a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3]};
        split8(v, hi, lo);
    } else {
        hi = *gl<bf16x8>((const __bf16*)p + off);
        lo = *gl<bf16x8>((const __bf16*)plo + off);
    }
}

__global__ __launch_bounds__(256) void k_gemm_mfma_sk(const GemmJob* jobs, const int32_t* bjob, const DesignDev* dd) {
    __shared__ float red[4][64][65];
    const int j = gl<int32_t>(bjob)[2 * blockIdx.x];
    if (j < 0) return;
    const int tb = gl<int32_t>(bjob)[2 * blockIdx.x + 1];
    const GemmJob& J = jobs[j];
    if (J.st->done) return;
    const DesignDev& D = dd[J.design];
    const int t = tb - J.tile_start;
    const int m0 = (t / J.tiles_n) * 64, n0 = (t % J.tiles_n) * 64;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const void* pa = J.a_kind ? J.A_hi : J.A;
    const void* pb = J.b_kind ? J.B_hi : J.B;
    f32x16 acc[4];
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
        for (int q = 0; q < 16; q++) acc[s][q] = 0.0f;
    const int kq = J.K / 4;
    const int kbeg = wv * kq, kend = kbeg + kq;
    const int r = lane & 31, kh = (lane >> 5) * 8;
    for (int k0 = kbeg; k0 < kend; k0 += 16) {
        bf16x8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
        for (int i = 0; i < 2; i++) {
            frag_direct(J.a_kind, pa, J.A_lo, m0 + 32 * i + r, k0 + kh, J.K, ah[i], al[i]);
            frag_direct(J.b_kind, pb, J.B_lo, n0 + 32 * i + r, k0 + kh, J.K, bh[i], bl[i]);
        }
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int jj = 0; jj < 2; jj++) {
                f32x16& c = acc[2 * i + jj];
                c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh[jj], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl[jj], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh[jj], c, 0, 0, 0);
            }
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int jj = 0; jj < 2; jj++)
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const int row = 32 * i + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
                const int col = 32 * jj + (lane & 31);
                red[wv][row][col] = acc[2 * i + jj][q];
            }
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * 64; e += 256) {
        const int row = e >> 6, col = e & 63;
        const float v = ((red[0][row][col] + red[1][row][col]) + red[2][row][col]) + red[3][row][col];
        epi_store<float>(J, D, m0 + row, n0 + col, (double)v);
    }
}

template <int TK, int NBUF, bool SUM>
static void launch_mfma(bool adata, const GemmJob* jobs, const int32_t* bjob, int total_tiles,
                        const DesignDev* dd, hipStream_t s) {
    if (adata)
        hipLaunchKernelGGL((k_gemm_mfma<TK, NBUF, true, SUM>), dim3(total_tiles), dim3(256), 0, s, jobs,
                           bjob, dd);
    else
        hipLaunchKernelGGL((k_gemm_mfma<TK, NBUF, false, SUM>), dim3(total_tiles), dim3(256), 0, s, jobs,
                           bjob, dd);
}

void launch_gemm_pass(const GemmJob* jobs, const int32_t* bjob, int total_tiles, const DesignDev* dd,
                      int prec, int engine, int pass, hipStream_t s) {
    if (total_tiles <= 0) return;
    // the data operand is A in F1 and I1 (passes 0, 2), B in F2 and I2
    const bool adata = (pass & 1) == 0;
    if (prec == EP_PREC_FP64)
        hipLaunchKernelGGL(k_gemm_fma<double>, dim3(total_tiles), dim3(256), 0, s, jobs, bjob, dd);
    else if (engine == 0)
        launch_mfma<32, 2, false>(adata, jobs, bjob, total_tiles, dd, s);
    else if (engine == 3)  // K stages of 64 through one LDS buffer (the LDS of engine 0)
        launch_mfma<64, 1, false>(adata, jobs, bjob, total_tiles, dd, s);
    else if (engine == 4 && pass > 0)  // K-sliced jobs reading split planes
        launch_mfma<32, 2, true>(adata, jobs, bjob, total_tiles, dd, s);
    else if (engine == 4)
        launch_mfma<32, 2, false>(adata, jobs, bjob, total_tiles, dd, s);
    else if (engine == 2)
        hipLaunchKernelGGL(k_gemm_mfma_sk, dim3(total_tiles), dim3(256), 0, s, jobs, bjob, dd);
    else
        hipLaunchKernelGGL(k_gemm_fma<float>, dim3(total_tiles), dim3(256), 0, s, jobs, bjob, dd);
}

}  // namespace ep
