// Spectral Poisson solve by fast transforms (density.rs:96-166), FAST path,
// bin counts that are powers of two.
//
// Three passes compute the field the four dense-product passes of
// ep_gemm.hip compute (Ex = C3_y (phi w_u) S3_x^T, Ey = S3_y (phi w_v) C3_x^T,
// phi the Poisson-solved DCT-II coefficients; the inverse may take y first
// because the transform is separable): a row pass transforms the density's
// rows (DCT-II along x) and stores them transposed; the column solve takes
// each of those rows through the DCT-II along y, the Poisson solve and both
// inverse transforms along y, and stores the two results transposed; a row
// pass per component finishes the inverse along x into the interleaved
// field, in row order.  Every 1-D transform of length n is one
// complex FFT of length n (Makhoul's reordering):
//   DCT-II  v = (x0, x2, x4, ..., x5, x3, x1), X_k = Re(e^{-i pi k/2n} V_k)
//   C3      W_k = e^{-i pi k/2n} (X_k + i X_{n-k}) with X_n = 0, w = FFT(W),
//           y_{2m} = Re(w_m) / 2, y_{2m+1} = Re(w_{n-1-m}) / 2
//   S3 (reference field)  y_j = (-1)^j C3(X reversed)_j
//   IDXST sine transform  y_j = -(-1)^j C3(0, X_{n-1}, ..., X_1)_j
// (sin(pi m (2j+1)/2n) = (-1)^j cos(pi (n-m)(2j+1)/2n) turns both sine
// transforms into C3).  The FFT of length n = N1 N2 takes two register
// phases (the four-step form): N2 threads per row each load the points
// t + N2 m (m < N1), take their N1-point DFT in registers and multiply by
// e^{-2 pi i t k1 / n}; one LDS exchange hands every thread N2 values of one
// k1, whose N2-point DFT gives the outputs k1 + N1 k2.  The outputs are
// staged in LDS and stored transposed in runs of consecutive rows.  Every
// value stays FP32; the twiddles are FP64 values rounded once.
#include <hip/hip_runtime.h>

#include <cmath>

#include "ep_internal.h"

namespace ep {

#define EP_GLOBAL __attribute__((address_space(1)))

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline f32x2 cmul(f32x2 a, f32x2 b) {
    return (f32x2){a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

// e^{-2 pi i k / 32} for k < 16 (cos, -sin)
__device__ inline f32x2 tw32(int k) {
    constexpr float C[16] = {1.0f, 0.98078528f, 0.923879533f, 0.831469612f, 0.707106781f,
                             0.555570233f, 0.382683432f, 0.195090322f, 0.0f, -0.195090322f,
                             -0.382683432f, -0.555570233f, -0.707106781f, -0.831469612f,
                             -0.923879533f, -0.98078528f};
    constexpr float S[16] = {0.0f, -0.195090322f, -0.382683432f, -0.555570233f, -0.707106781f,
                             -0.831469612f, -0.923879533f, -0.98078528f, -1.0f, -0.98078528f,
                             -0.923879533f, -0.831469612f, -0.707106781f, -0.555570233f,
                             -0.382683432f, -0.195090322f};
    return (f32x2){C[k], S[k]};
}

template <int N>
__device__ constexpr int brev(int i) {
    int r = 0;
    for (int b = 1; b < N; b <<= 1) {
        r = (r << 1) | (i & 1);
        i >>= 1;
    }
    return r;
}

// In-register DFT of N <= 32 points (radix-2, decimation in time; every
// index is a compile-time constant once the loops are unrolled).
template <int N>
__device__ inline void dft_reg(f32x2 (&a)[N]) {
    f32x2 b[N];
#pragma unroll
    for (int i = 0; i < N; i++) b[brev<N>(i)] = a[i];
#pragma unroll
    for (int len = 2; len <= N; len <<= 1) {
#pragma unroll
        for (int i = 0; i < N; i += len) {
#pragma unroll
            for (int j = 0; j < len / 2; j++) {
                f32x2 v = b[i + j + len / 2];
                if (j != 0) {
                    if (4 * j == len) v = (f32x2){v.y, -v.x};  // times -i
                    else v = cmul(v, tw32(j * (32 / len)));
                }
                const f32x2 u = b[i + j];
                b[i + j] = u + v;
                b[i + j + len / 2] = u - v;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < N; i++) a[i] = b[i];
}

// Split of each supported length n = N1 N2 (N1 points per thread in the
// first phase, N2 threads per row; a block of kFftThreads threads takes kFftThreads / N2
// rows).  Row passes: 16 = 4 x 4, 32 = 8 x 4, 64 = 8 x 8, 128 = 16 x 8,
// 256 = 16 x 16, 512 = 16 x 32.  The column solve keeps more in LDS per row
// and takes 128 as 8 x 16.
__host__ __device__ constexpr int fft_n2(int n) {
    return n >= 512 ? 32 : n >= 256 ? 16 : n >= 64 ? 8 : 4;
}
__host__ __device__ constexpr int fft_n2_solve(int n) {
    return n >= 512 ? 32 : n >= 128 ? 16 : n >= 64 ? 8 : 4;
}
// Threads per block of the FFT kernels.
constexpr int kFftThreads = 256;
__host__ __device__ constexpr int fft_block_rows(int n) { return kFftThreads / fft_n2(n); }
__host__ __device__ constexpr int fft_solve_rows(int n) { return kFftThreads / fft_n2_solve(n); }

// LDS of the row pass (largest case: n = 128, kFftThreads / 8 rows x 16 x
// (8 + 1) points)
constexpr int kFftLds = (kFftThreads / 8) * 16 * 9;  // f32x2 elements
// LDS of the column solve: the exchange (largest: n = 256, kFftThreads / 16
// rows x 16 x 17 points) and the Poisson-solved rows (largest: n = 256,
// kFftThreads / 16 rows x 257)
constexpr int kSolveEx = (kFftThreads / 16) * 16 * 17;  // f32x2 elements
constexpr int kSolveRows = (kFftThreads / 16) * 257;    // floats

// Transform families of a call site, known when it is compiled: the DCT-II,
// C3, and the two sine kinds (DST-III of the reference field, the IDXST
// transform), between which the job's kind chooses at run time.
enum FftFamily { kFamDct2 = 0, kFamC3 = 1, kFamSine = 2 };

// The complex input points j = t + N2 m (m < N1) of thread t's share of one
// row: DCT-II takes v_j = x_{2j} (j < n/2) or x_{2(n-1-j)+1}; the C3 kinds
// take e^{-i pi j/2n} (Z_j + i Z_{n-j}), Z_n = 0, with Z the row (C3), the
// reversed row (DST-III, kind 2) or (0, x_{n-1}, ..., x_1) (IDXST, kind 3).
// row(i) returns element i of the row.  The loads of a family are the same
// whatever the run-time kind (indices wrapped into the row where the element
// is zeroed), so no load's address depends on a branch.  The factor
// e^{-i pi j/2n} is e^{-i pi t/2n} (one load per thread) times
// e^{-i pi N2 m/2n} (the same address in every lane: a scalar load).
template <int N1, int N2, int F, typename Row>
__device__ inline void fft_input(int kind, int t, const EP_GLOBAL f32x2* dtw, Row row, f32x2 (&a)[N1]) {
    constexpr int n = N1 * N2;
    const f32x2 dt = F == kFamDct2 ? (f32x2){1.f, 0.f} : dtw[t];
#pragma unroll
    for (int m = 0; m < N1; m++) {
        const int j = t + N2 * m;
        if (F == kFamDct2) {
            a[m] = (f32x2){row(j < n / 2 ? 2 * j : 2 * (n - 1 - j) + 1), 0.f};
        } else {
            const int ia = F == kFamC3 ? j : kind == 2 ? n - 1 - j : (n - j) & (n - 1);
            const int ib = F == kFamC3 ? (n - j) & (n - 1) : kind == 2 ? (j - 1) & (n - 1) : j;
            const float xa = row(ia), xb = row(ib);
            const f32x2 c = m == 0 ? dt : cmul(dt, dtw[N2 * m]);
            const float za = (F == kFamSine && kind == 3 && j == 0) ? 0.f : xa;
            const float zb = j == 0 ? 0.f : xb;
            a[m] = cmul(c, (f32x2){za, zb});
        }
    }
}

// Phase 1: the N1-point DFT of the thread's points, times e^{-2 pi i t k1/n},
// into the exchange (row r, k1, t).  The twiddle table is laid out for the
// split, e^{-2 pi i t k1/n} at k1 N2 + t, so for each k1 the lanes of a wave
// read consecutive entries (one or two cache lines, where t k1 order would
// touch one line per lane).
template <int N1, int N2>
__device__ inline void fft_phase1(f32x2 (&a)[N1], const EP_GLOBAL f32x2* tw, f32x2* ex, int r, int t) {
    constexpr int PAD = N2 + 1;
    dft_reg<N1>(a);
#pragma unroll
    for (int k1 = 0; k1 < N1; k1++) ex[(r * N1 + k1) * PAD + t] = cmul(a[k1], tw[k1 * N2 + t]);
}

// The thread's phase-1 twiddles in registers, for kernels that take several
// transforms of the same length (loaded once, with the kernel's first loads).
template <int N1, int N2>
__device__ inline void load_tw(const EP_GLOBAL f32x2* tw, int t, f32x2 (&w)[N1]) {
#pragma unroll
    for (int k1 = 0; k1 < N1; k1++) w[k1] = tw[k1 * N2 + t];
}
template <int N1, int N2>
__device__ inline void fft_phase1(f32x2 (&a)[N1], const f32x2 (&w)[N1], f32x2* ex, int r, int t) {
    constexpr int PAD = N2 + 1;
    dft_reg<N1>(a);
#pragma unroll
    for (int k1 = 0; k1 < N1; k1++) ex[(r * N1 + k1) * PAD + t] = cmul(a[k1], w[k1]);
}

// Output index of phase-2 value (q, k2) of thread t: k = k1 + N1 k2 with
// k1 = t + N2 q; DCT-II writes it at k, the C3 kinds at j = 2k (k < n/2) or
// 2(n-1-k)+1.
template <int N1, int N2, int F>
__device__ inline int fft_out_index(int t, int q, int k2) {
    constexpr int n = N1 * N2;
    const int k = t + N2 * q + N1 * k2;
    return F == kFamDct2 ? k : (k < n / 2 ? 2 * k : 2 * (n - 1 - k) + 1);
}

// Phase 2: for this thread's k1 = t + N2 q (< N1), the N2-point DFT of the
// exchanged values; emit(q, k2, j, value, c) receives output k = k1 + N1 k2 of
// the real transform at its index j: DCT-II X_k = Re(c V_k) with
// c = e^{-i pi k/2n} (passed on); C3 kinds Re(w_k) / 2, signed (-1)^j for
// DST-III and -(-1)^j for IDXST (c = 0).
template <int N1, int N2, int F, typename Emit>
__device__ inline void fft_phase2(const f32x2* ex, int r, int t, int kind, const EP_GLOBAL f32x2* dtw,
                                  Emit emit) {
    constexpr int PAD = N2 + 1, NQ = (N1 + N2 - 1) / N2;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const int k1 = t + N2 * q;
        if (k1 < N1) {
            f32x2 b[N2];
#pragma unroll
            for (int s2 = 0; s2 < N2; s2++) b[s2] = ex[(r * N1 + k1) * PAD + s2];
            // DCT-II: e^{-i pi k/2n} = e^{-i pi k1/2n} (one load) e^{-i pi N1 k2/2n} (a scalar load)
            const f32x2 dk = F == kFamDct2 ? dtw[k1] : (f32x2){1.f, 0.f};
            dft_reg<N2>(b);
#pragma unroll
            for (int k2 = 0; k2 < N2; k2++) {
                const int j = fft_out_index<N1, N2, F>(t, q, k2);
                if (F == kFamDct2) {
                    const f32x2 c = k2 == 0 ? dk : cmul(dk, dtw[N1 * k2]);
                    emit(q, k2, j, c.x * b[k2].x - c.y * b[k2].y, c);
                } else {
                    const float y = 0.5f * b[k2].x;
                    const bool neg = F == kFamSine && ((kind == 2 && (j & 1)) || (kind == 3 && !(j & 1)));
                    emit(q, k2, j, neg ? -y : y, (f32x2){0.f, 0.f});
                }
            }
        }
    }
}

// Values kept in registers between phase 2 and the staging, one per output.
template <int N1, int N2>
struct FftOut {
    float v[((N1 + N2 - 1) / N2) * N2];
};

// The thread's outputs into a staged row (stride n + 1 floats).
template <int N1, int N2, int F>
__device__ inline void fft_stage(float* stg, int r, int t, const FftOut<N1, N2>& o) {
    constexpr int n = N1 * N2, NQ = (N1 + N2 - 1) / N2;
#pragma unroll
    for (int q = 0; q < NQ; q++)
        if (t + N2 * q < N1) {
#pragma unroll
            for (int k2 = 0; k2 < N2; k2++)
                stg[r * (n + 1) + fft_out_index<N1, N2, F>(t, q, k2)] = o.v[q * N2 + k2];
        }
}

// Store of rh staged rows of n: transposed (element k of row rr to
// out[(k ld + r0 + rr) os], consecutive lanes on consecutive rows) or in
// row order (out[((r0 + rr) n + k) os], consecutive lanes on consecutive k).
template <int n, int R, bool ROWS>
__device__ inline void fft_store(const float* stg, EP_GLOBAL float* out, int ld, int os, int r0, int rh) {
    const int tid = threadIdx.x;
    auto put = [&](int rr, int k) {
        const size_t o = ROWS ? (size_t)(r0 + rr) * n + k : (size_t)k * ld + r0 + rr;
        out[o * os] = stg[rr * (n + 1) + k];
    };
    if (rh == R) {
#pragma unroll
        for (int u = 0; u < R * n / kFftThreads; u++) {
            const int i = tid + kFftThreads * u;
            if (ROWS) put(i / n, i % n);
            else put(i % R, i / R);
        }
    } else {
        const int lrh = 31 - __clz(rh);
        for (int i = tid; i < rh * n; i += kFftThreads) {
            if (ROWS) put(i / n, i % n);
            else put(i & (rh - 1), i >> lrh);
        }
    }
}

// The first pass: rows r0 .. r0 + rh of the density, DCT-II along x, stored
// transposed (element k of row r at out[k ld + r]).
template <int N1, int N2>
__device__ inline void fft_rows(const FftJob& J, int r0, f32x2* ex) {
    constexpr int n = N1 * N2;
    constexpr int R = kFftThreads / N2;
    static_assert(R * N1 * (N2 + 1) <= kFftLds && R * (n + 1) <= 2 * kFftLds, "LDS");
    const int tid = threadIdx.x;
    const int t = tid % N2, r = tid / N2;
    const int rh = min(R, J.rows - r0);  // rows of this block (a power of two)
    const bool live = r < rh;
    const EP_GLOBAL float* x = (const EP_GLOBAL float*)J.in + (size_t)(r0 + r) * n;
    const EP_GLOBAL f32x2* tw = (const EP_GLOBAL f32x2*)J.tw;
    const EP_GLOBAL f32x2* dtw = (const EP_GLOBAL f32x2*)J.dtw;
    if (live) {
        f32x2 a[N1];
        fft_input<N1, N2, kFamDct2>(0, t, dtw, [&](int i) { return x[i]; }, a);
        fft_phase1<N1, N2>(a, tw, ex, r, t);
    }
    __syncthreads();
    FftOut<N1, N2> o;
    if (live)
        fft_phase2<N1, N2, kFamDct2>(ex, r, t, 0, dtw, [&](int q, int k2, int, float v, f32x2) { o.v[q * N2 + k2] = v; });
    __syncthreads();  // the exchange is read; its LDS now stages the outputs
    float* stg = (float*)ex;
    if (live) fft_stage<N1, N2, kFamDct2>(stg, r, t, o);
    __syncthreads();
    fft_store<n, R, false>(stg, (EP_GLOBAL float*)J.out, J.ld, 1, r0, rh);
}

// The column solve (density.rs:113-153 between the two transforms): rows u
// of the plane transformed along x, each taken through the DCT-II along y,
// the Poisson solve phi = a s / (l_u + l_v) (s = 4 / (Bx By), halved at u = 0
// and at v = 0; zero where the denominator vanishes), and both inverse
// transforms along y: C3 of phi w_u (the Ex path) and the job's sine kind of
// phi w_v (the Ey path), stored transposed into out (Ex) and out2 (Ey).  The
// inverse along x then finishes each component in a row pass.
//
// The factors along y are not loaded: l_v = 2 (1 - cos(pi v / n)) =
// 4 sin^2(pi v / 2n) comes from the DCT-II factor of the same output,
// e^{-i pi v/2n}, which phase 2 loads anyway (within two ulp of the table
// value), and w_v = pi v / n is computed as the upload computes it (n a power
// of two, so the quotient is exact and the value the same bit for bit).  The
// quotient takes the hardware reciprocal (1 ulp): a division per coefficient
// doubled the solve's time.
template <int N1, int N2>
__device__ inline void fft_solve(const FftJob& J, const DesignDev* dd, int r0, f32x2* ex, float* phi) {
    constexpr int n = N1 * N2;
    constexpr int R = kFftThreads / N2;
    static_assert(R * N1 * (N2 + 1) <= kSolveEx && R * (n + 1) <= kSolveRows &&
                      R * (n + 1) <= 2 * kSolveEx, "LDS");
    const int tid = threadIdx.x;
    const int t = tid % N2, r = tid / N2;
    const int rh = min(R, J.rows - r0);
    const bool live = r < rh;
    const int kx = r0 + r;
    const EP_GLOBAL float* x = (const EP_GLOBAL float*)J.in + (size_t)kx * n;
    const EP_GLOBAL f32x2* tw = (const EP_GLOBAL f32x2*)J.tw;
    const EP_GLOBAL f32x2* dtw = (const EP_GLOBAL f32x2*)J.dtw;
    const DesignDev& D = dd[J.design];
    const EP_GLOBAL f32x2* sx = (const EP_GLOBAL f32x2*)D.spec_x;
    const float s0 = 4.0f / ((float)D.Bx * (float)D.By);
    float* prow = phi + r * (n + 1);
    f32x2 twr[N1];  // the three transforms' phase-1 twiddles
    // forward DCT-II along y, the Poisson solve written into the LDS rows
    if (live) {
        f32x2 a[N1];
        load_tw<N1, N2>(tw, t, twr);
        fft_input<N1, N2, kFamDct2>(0, t, dtw, [&](int i) { return x[i]; }, a);
        fft_phase1<N1, N2>(a, twr, ex, r, t);
    }
    __syncthreads();
    if (live) {
        const f32x2 fx = sx[kx];
        const float su = kx == 0 ? 0.5f * s0 : s0;
        fft_phase2<N1, N2, kFamDct2>(ex, r, t, 0, dtw, [&](int, int, int k, float v, f32x2 c) {
            const float lv = 4.f * (c.y * c.y);
            const float sc = k == 0 ? 0.5f * su : su;
            const float den = fx.y + lv;
            prow[k] = den < 1e-30f ? 0.f : v * sc * __builtin_amdgcn_rcpf(den);
        });
    }
    __syncthreads();
    // Ex path: C3 along y of phi w_u
    FftOut<N1, N2> oa;
    if (live) {
        const float wu = sx[kx].x;
        f32x2 a[N1];
        fft_input<N1, N2, kFamC3>(1, t, dtw, [&](int i) { return prow[i] * wu; }, a);
        fft_phase1<N1, N2>(a, twr, ex, r, t);
    }
    __syncthreads();
    if (live)
        fft_phase2<N1, N2, kFamC3>(ex, r, t, 1, dtw, [&](int q, int k2, int, float v, f32x2) { oa.v[q * N2 + k2] = v; });
    __syncthreads();
    // Ey path: the sine kind along y of phi w_v
    const int kind = J.kind;
    if (live) {
        f32x2 a[N1];
        constexpr double kPi = 3.14159265358979323846;
        fft_input<N1, N2, kFamSine>(kind, t, dtw, [&](int i) { return prow[i] * (float)(kPi * (double)i / n); },
                                    a);
        fft_phase1<N1, N2>(a, twr, ex, r, t);
    }
    __syncthreads();  // the Poisson rows are read: they now stage the Ex output
    float* stA = phi;
    if (live) fft_stage<N1, N2, kFamC3>(stA, r, t, oa);
    FftOut<N1, N2> ob;
    if (live)
        fft_phase2<N1, N2, kFamSine>(ex, r, t, kind, dtw, [&](int q, int k2, int, float v, f32x2) { ob.v[q * N2 + k2] = v; });
    __syncthreads();  // the exchange is read; its LDS now stages the Ey output
    float* stB = (float*)ex;
    if (live) fft_stage<N1, N2, kFamSine>(stB, r, t, ob);
    __syncthreads();
    fft_store<n, R, false>(stA, (EP_GLOBAL float*)J.out, J.ld, 1, r0, rh);
    fft_store<n, R, false>(stB, (EP_GLOBAL float*)J.out2, J.ld, 1, r0, rh);
}

// The last pass for both components at once: rows r0 .. r0 + rh of plane in
// (the job's sine kind along x, Ex) and of plane in2 (C3 along x, Ey), stored
// in row order as (Ex, Ey) pairs into out.  The second plane's points are
// loaded before the first plane's exchange is read, so their latency
// overlaps it.
template <int N1, int N2>
__device__ inline void fft_pair(const FftJob& J, int r0, f32x2* ex) {
    constexpr int n = N1 * N2;
    constexpr int R = kFftThreads / N2;
    static_assert(R * N1 * (N2 + 1) <= kFftLds && 2 * R * (n + 1) <= 2 * kFftLds, "LDS");
    const int tid = threadIdx.x;
    const int t = tid % N2, r = tid / N2;
    const int rh = min(R, J.rows - r0);
    const bool live = r < rh;
    const int kind = J.kind;
    const EP_GLOBAL float* xa = (const EP_GLOBAL float*)J.in + (size_t)(r0 + r) * n;
    const EP_GLOBAL float* xb = (const EP_GLOBAL float*)J.in2 + (size_t)(r0 + r) * n;
    const EP_GLOBAL f32x2* tw = (const EP_GLOBAL f32x2*)J.tw;
    const EP_GLOBAL f32x2* dtw = (const EP_GLOBAL f32x2*)J.dtw;
    f32x2 b[N1], twr[N1];  // twr: both transforms' phase-1 twiddles
    if (live) {
        f32x2 a[N1];
        load_tw<N1, N2>(tw, t, twr);
        fft_input<N1, N2, kFamSine>(kind, t, dtw, [&](int i) { return xa[i]; }, a);
        fft_input<N1, N2, kFamC3>(1, t, dtw, [&](int i) { return xb[i]; }, b);
        fft_phase1<N1, N2>(a, twr, ex, r, t);
    }
    __syncthreads();
    FftOut<N1, N2> oa, ob;
    if (live)
        fft_phase2<N1, N2, kFamSine>(ex, r, t, kind, dtw, [&](int q, int k2, int, float v, f32x2) { oa.v[q * N2 + k2] = v; });
    __syncthreads();
    if (live) fft_phase1<N1, N2>(b, twr, ex, r, t);
    __syncthreads();
    if (live)
        fft_phase2<N1, N2, kFamC3>(ex, r, t, 1, dtw, [&](int q, int k2, int, float v, f32x2) { ob.v[q * N2 + k2] = v; });
    __syncthreads();  // the exchange is read; its LDS now stages (Ex, Ey) pairs
    f32x2* stg = ex;
    if (live) {
        constexpr int NQ = (N1 + N2 - 1) / N2;
#pragma unroll
        for (int q = 0; q < NQ; q++)
            if (t + N2 * q < N1) {
#pragma unroll
                for (int k2 = 0; k2 < N2; k2++) {
                    const int j = fft_out_index<N1, N2, kFamC3>(t, q, k2);
                    stg[r * (n + 1) + j] = (f32x2){oa.v[q * N2 + k2], ob.v[q * N2 + k2]};
                }
            }
    }
    __syncthreads();
    EP_GLOBAL f32x2* out = (EP_GLOBAL f32x2*)J.out;
    if (rh == R) {
#pragma unroll
        for (int u = 0; u < R * n / kFftThreads; u++) {
            const int i = tid + kFftThreads * u;
            const int rr = i / n, k = i % n;
            out[(size_t)(r0 + rr) * n + k] = stg[rr * (n + 1) + k];
        }
    } else {
        for (int i = tid; i < rh * n; i += kFftThreads) {
            const int rr = i / n, k = i % n;
            out[(size_t)(r0 + rr) * n + k] = stg[rr * (n + 1) + k];
        }
    }
}

__global__ __launch_bounds__(kFftThreads) void k_fft_pair(const FftJob* jobs, const int32_t* fblk) {
    __shared__ f32x2 ex[kFftLds];
    const int jb = ((const EP_GLOBAL int32_t*)fblk)[2 * blockIdx.x];
    if (jb < 0) return;
    const int r0 = ((const EP_GLOBAL int32_t*)fblk)[2 * blockIdx.x + 1];
    const FftJob& J = jobs[jb];
    if (J.st->done) return;
    switch (J.n) {
        case 16: fft_pair<4, 4>(J, r0, ex); break;
        case 32: fft_pair<8, 4>(J, r0, ex); break;
        case 64: fft_pair<8, 8>(J, r0, ex); break;
        case 128: fft_pair<16, 8>(J, r0, ex); break;
        case 256: fft_pair<16, 16>(J, r0, ex); break;
        case 512: fft_pair<16, 32>(J, r0, ex); break;
        default: break;
    }
}

__global__ __launch_bounds__(kFftThreads) void k_fft_pass(const FftJob* jobs, const int32_t* fblk) {
    __shared__ f32x2 ex[kFftLds];
    const int jb = ((const EP_GLOBAL int32_t*)fblk)[2 * blockIdx.x];
    if (jb < 0) return;
    const int r0 = ((const EP_GLOBAL int32_t*)fblk)[2 * blockIdx.x + 1];
    const FftJob& J = jobs[jb];
    if (J.st->done) return;
    switch (J.n) {
        case 16: fft_rows<4, 4>(J, r0, ex); break;
        case 32: fft_rows<8, 4>(J, r0, ex); break;
        case 64: fft_rows<8, 8>(J, r0, ex); break;
        case 128: fft_rows<16, 8>(J, r0, ex); break;
        case 256: fft_rows<16, 16>(J, r0, ex); break;
        case 512: fft_rows<16, 32>(J, r0, ex); break;
        default: break;
    }
}

__global__ __launch_bounds__(kFftThreads) void k_fft_solve(const FftJob* jobs, const int32_t* fblk,
                                                   const DesignDev* dd) {
    __shared__ f32x2 ex[kSolveEx];
    __shared__ float phi[kSolveRows];
    const int jb = ((const EP_GLOBAL int32_t*)fblk)[2 * blockIdx.x];
    if (jb < 0) return;
    const int r0 = ((const EP_GLOBAL int32_t*)fblk)[2 * blockIdx.x + 1];
    const FftJob& J = jobs[jb];
    if (J.st->done) return;
    switch (J.n) {
        case 16: fft_solve<4, 4>(J, dd, r0, ex, phi); break;
        case 32: fft_solve<8, 4>(J, dd, r0, ex, phi); break;
        case 64: fft_solve<8, 8>(J, dd, r0, ex, phi); break;
        case 128: fft_solve<8, 16>(J, dd, r0, ex, phi); break;
        case 256: fft_solve<16, 16>(J, dd, r0, ex, phi); break;
        case 512: fft_solve<16, 32>(J, dd, r0, ex, phi); break;
        default: break;
    }
}

int fft_rows_per_block(int n) { return fft_block_rows(n); }
int fft_solve_rows_per_block(int n) { return fft_solve_rows(n); }

void fft_twiddle_table(int n, bool solve, float* out) {
    const int n2 = solve ? fft_n2_solve(n) : fft_n2(n), n1 = n / n2;
    const double pi = 3.14159265358979323846;
    for (int k1 = 0; k1 < n1; k1++)
        for (int t = 0; t < n2; t++) {
            const double a = 2.0 * pi * (double)(t * k1) / n;
            out[2 * (k1 * n2 + t)] = (float)std::cos(a);
            out[2 * (k1 * n2 + t) + 1] = (float)-std::sin(a);
        }
}

void launch_fft_pass(const FftJob* jobs, const int32_t* fblk, int blocks, const DesignDev* dd,
                     hipStream_t s) {
    if (blocks <= 0) return;
    hipLaunchKernelGGL(k_fft_pass, dim3(blocks), dim3(kFftThreads), 0, s, jobs, fblk);
}

void launch_fft_pair(const FftJob* jobs, const int32_t* fblk, int blocks, hipStream_t s) {
    if (blocks <= 0) return;
    hipLaunchKernelGGL(k_fft_pair, dim3(blocks), dim3(kFftThreads), 0, s, jobs, fblk);
}

void launch_fft_solve(const FftJob* jobs, const int32_t* fblk, int blocks, const DesignDev* dd,
                      hipStream_t s) {
    if (blocks <= 0) return;
    hipLaunchKernelGGL(k_fft_solve, dim3(blocks), dim3(kFftThreads), 0, s, jobs, fblk, dd);
}

}  // namespace ep
