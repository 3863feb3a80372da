// Spectral Poisson solve by fast transforms (density.rs:96-166), FAST path,
// bin counts that are powers of two.
//
// The passes compute what the dense-product passes of ep_gemm.hip compute:
// F1 transforms the density's rows (DCT-II along x) and stores them
// transposed, F2 transforms those rows (DCT-II along y) and applies the
// Poisson solve, I1 transforms the rows of Ex^ and Ey^ (S3 and C3 along x)
// and stores them transposed, I2 transforms those rows (C3 and S3 along y)
// into the interleaved field.  Every 1-D transform of length n is one
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
// first phase, N2 threads per row): 16 = 4 x 4, 32 = 8 x 4, 64 = 8 x 8,
// 128 = 16 x 8, 256 = 16 x 16, 512 = 16 x 32.  A block of 256 threads takes
// 256 / N2 rows.
__host__ __device__ constexpr int fft_n2(int n) {
    return n >= 512 ? 32 : n >= 256 ? 16 : n >= 64 ? 8 : 4;
}
__host__ __device__ constexpr int fft_block_rows(int n) { return 256 / fft_n2(n); }

// LDS of the largest case (n = 128: 32 rows x 16 x (8 + 1) points)
constexpr int kFftLds = 32 * 16 * 9;  // f32x2 elements

template <int N1, int N2>
__device__ inline void fft_rows(const FftJob& J, const DesignDev* dd, int r0, f32x2* ex) {
    constexpr int n = N1 * N2;
    constexpr int R = 256 / N2;
    constexpr int PAD = N2 + 1;
    constexpr int NQ = (N1 + N2 - 1) / N2;  // second-phase DFTs per thread (some idle when N1 < N2)
    constexpr int NO = NQ * N2;              // outputs per thread of the second phase
    static_assert(R * N1 * PAD <= kFftLds && R * (n + 1) <= 2 * kFftLds, "LDS");
    const int tid = threadIdx.x;
    const int t = tid % N2, r = tid / N2;
    const int rh = min(R, J.rows - r0);  // rows of this block (a power of two)
    const bool live = r < rh;
    const int kind = J.kind;
    const EP_GLOBAL float* x = (const EP_GLOBAL float*)J.in + (size_t)(r0 + r) * n;
    const EP_GLOBAL f32x2* tw = (const EP_GLOBAL f32x2*)J.tw;
    const EP_GLOBAL f32x2* dtw = (const EP_GLOBAL f32x2*)J.dtw;

    // phase 1: points j = t + N2 m, scaled by e^{-i pi j/2n} for the C3
    // kinds; N1-point DFT; times e^{-2 pi i t k1/n}.  The loads are the
    // same for every kind (two in-row elements and the factor; the kind only
    // picks indices and selects): the address of a load never depends on a
    // branch over the kind.
    if (live) {
        f32x2 a[N1];
#pragma unroll
        for (int m = 0; m < N1; m++) {
            const int j = t + N2 * m;
            // DCT-II: v_j = x_{2j} (j < n/2) or x_{2(n-1-j)+1}; C3 kinds:
            // Z_j and Z_{n-j} (Z_n = 0), indices wrapped into the row where the
            // element is zeroed below
            const int i0 = j < n / 2 ? 2 * j : 2 * (n - 1 - j) + 1;
            const int ia = kind == 0 ? i0 : kind == 1 ? j : kind == 2 ? n - 1 - j : (n - j) & (n - 1);
            const int ib = kind == 1 ? (n - j) & (n - 1) : kind == 2 ? (j - 1) & (n - 1) : j;
            const float xa = x[ia], xb = x[ib];
            const f32x2 c = dtw[j];
            const float za = (kind == 3 && j == 0) ? 0.f : xa;
            const float zb = j == 0 ? 0.f : xb;
            const f32x2 w = cmul(c, (f32x2){za, zb});
            a[m] = kind == 0 ? (f32x2){xa, 0.f} : w;
        }
        dft_reg<N1>(a);
#pragma unroll
        for (int k1 = 0; k1 < N1; k1++) ex[(r * N1 + k1) * PAD + t] = cmul(a[k1], tw[t * k1]);
    }
    __syncthreads();
    // phase 2: this thread's k1 = t + N2 q (< N1); outputs k = k1 + N1 k2
    float val[NO];
    int oj[NO];
    if (live) {
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            const int k1 = t + N2 * q;
            if (k1 < N1) {
                f32x2 b[N2];
#pragma unroll
                for (int s2 = 0; s2 < N2; s2++) b[s2] = ex[(r * N1 + k1) * PAD + s2];
                dft_reg<N2>(b);
#pragma unroll
                for (int k2 = 0; k2 < N2; k2++) {
                    // DCT-II: X_k = Re(e^{-i pi k/2n} V_k) at k; C3 kinds:
                    // Re(w_k) / 2 at j (signed for the sine kinds)
                    const int u = q * N2 + k2, k = k1 + N1 * k2;
                    const f32x2 c = dtw[k];
                    const int j = k < n / 2 ? 2 * k : 2 * (n - 1 - k) + 1;
                    const float x0 = c.x * b[k2].x - c.y * b[k2].y;
                    const float y = 0.5f * b[k2].x;
                    const bool neg = (kind == 2 && (j & 1)) || (kind == 3 && !(j & 1));
                    oj[u] = kind == 0 ? k : j;
                    val[u] = kind == 0 ? x0 : neg ? -y : y;
                }
            }
        }
    }
    __syncthreads();  // the exchange is read; its LDS now stages the outputs
    float* stg = (float*)ex;
    if (live) {
#pragma unroll
        for (int q = 0; q < NQ; q++)
            if (t + N2 * q < N1) {
#pragma unroll
                for (int k2 = 0; k2 < N2; k2++) stg[r * (n + 1) + oj[q * N2 + k2]] = val[q * N2 + k2];
            }
    }
    __syncthreads();
    // transposed store: consecutive lanes take consecutive rows of one output
    // index; a full block stores R n / 256 rounds of 256 elements
    const int ld = J.ld;
    EP_GLOBAL float* out = (EP_GLOBAL float*)J.out;
    if (J.epi == 1) {
        // the Poisson solve (density.rs:113-153) at (ky = k, kx = r0 + rr)
        const DesignDev& D = dd[J.design];
        const EP_GLOBAL f32x2* sx = (const EP_GLOBAL f32x2*)D.spec_x;
        const EP_GLOBAL f32x2* sy = (const EP_GLOBAL f32x2*)D.spec_y;
        EP_GLOBAL float* out2 = (EP_GLOBAL float*)J.out2;
        const float s0 = 4.0f / ((float)D.Bx * (float)D.By);
        auto poisson_store = [&](int rr, int k) {
            const int kx = r0 + rr;
            const f32x2 fx = sx[kx], fy = sy[k];
            float sc = s0;
            if (kx == 0) sc *= 0.5f;
            if (k == 0) sc *= 0.5f;
            const float a = stg[rr * (n + 1) + k] * sc;
            const float den = fx.y + fy.y;
            float e0 = 0.f, e1 = 0.f;
            if (!(den < 1e-30f)) {
                const float phi = a / den;
                e0 = phi * fx.x;
                e1 = phi * fy.x;
            }
            const size_t o = (size_t)k * ld + kx;
            out[o] = e0;
            out2[o] = e1;
        };
        if (rh == R) {
#pragma unroll
            for (int u = 0; u < R * n / 256; u++) {
                const int i = tid + 256 * u;
                poisson_store(i % R, i / R);
            }
        } else {
            const int lrh = 31 - __clz(rh);
            for (int i = tid; i < rh * n; i += 256) poisson_store(i & (rh - 1), i >> lrh);
        }
        return;
    }
    const int os = J.ostride;
    if (rh == R) {
#pragma unroll
        for (int u = 0; u < R * n / 256; u++) {
            const int i = tid + 256 * u;
            const int rr = i % R, k = i / R;
            out[((size_t)k * ld + r0 + rr) * os] = stg[rr * (n + 1) + k];
        }
    } else {
        const int lrh = 31 - __clz(rh);
        for (int i = tid; i < rh * n; i += 256) {
            const int rr = i & (rh - 1), k = i >> lrh;
            out[((size_t)k * ld + r0 + rr) * os] = stg[rr * (n + 1) + k];
        }
    }
}

__global__ __launch_bounds__(256) void k_fft_pass(const FftJob* jobs, const int32_t* fblk,
                                                  const DesignDev* dd) {
    __shared__ f32x2 ex[kFftLds];
    const int jb = ((const EP_GLOBAL int32_t*)fblk)[2 * blockIdx.x];
    if (jb < 0) return;
    const int r0 = ((const EP_GLOBAL int32_t*)fblk)[2 * blockIdx.x + 1];
    const FftJob& J = jobs[jb];
    if (J.st->done) return;
    switch (J.n) {
        case 16: fft_rows<4, 4>(J, dd, r0, ex); break;
        case 32: fft_rows<8, 4>(J, dd, r0, ex); break;
        case 64: fft_rows<8, 8>(J, dd, r0, ex); break;
        case 128: fft_rows<16, 8>(J, dd, r0, ex); break;
        case 256: fft_rows<16, 16>(J, dd, r0, ex); break;
        case 512: fft_rows<16, 32>(J, dd, r0, ex); break;
        default: break;
    }
}

int fft_rows_per_block(int n) { return fft_block_rows(n); }

void launch_fft_pass(const FftJob* jobs, const int32_t* fblk, int blocks, const DesignDev* dd,
                     hipStream_t s) {
    if (blocks <= 0) return;
    hipLaunchKernelGGL(k_fft_pass, dim3(blocks), dim3(256), 0, s, jobs, fblk, dd);
}

}  // namespace ep
