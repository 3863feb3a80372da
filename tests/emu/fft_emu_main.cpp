// CPU build of the FFT pass kernel of eisenjoch_amd/csrc/ep_fft.hip: the
// kernel's own source (extracted into fft_kernel_extract.inc by the test) runs
// with 256 std::threads per block and a std::barrier for __syncthreads, on
// exact-size heap buffers (AddressSanitizer reports any access outside them),
// and its outputs are checked against direct O(n^2) transforms.  Arguments:
// the row lengths to run (default: every supported length).
#include <algorithm>
#include <barrier>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>
#define __device__
#define __host__
#define __global__
#define __shared__ static
#define __launch_bounds__(...)
#define EP_GLOBAL
using std::min;
struct Dim { int x; };
static thread_local Dim threadIdx;
static Dim blockIdx;
static std::barrier<>* g_bar;
#define __syncthreads() g_bar->arrive_and_wait()
#define __clz(v) __builtin_clz(v)
struct DevState { int done; };
struct DesignDev { int Bx, By; const void* spec_x; const void* spec_y; };
struct FftJob {
    const void* in; void* out; void* out2; const void* tw; const void* dtw; const DevState* st;
    int32_t n, rows, rpb, kind, epi, ld, ostride, design;
};
#include "fft_kernel_extract.inc"

static void run_pass(const std::vector<FftJob>& jobs, const std::vector<int32_t>& fblk, const DesignDev* dd) {
    for (size_t b = 0; b < fblk.size() / 2; b++) {
        blockIdx.x = (int)b;
        std::barrier<> bar(256);
        g_bar = &bar;
        std::vector<std::thread> th;
        for (int t = 0; t < 256; t++)
            th.emplace_back([&, t] { threadIdx.x = t; k_fft_pass(jobs.data(), fblk.data(), dd); });
        for (auto& x : th) x.join();
    }
}

int main(int argc, char** argv) {
    std::vector<int> lens;
    for (int a = 1; a < argc; a++) lens.push_back(atoi(argv[a]));
    if (lens.empty()) lens = {16, 32, 64, 128, 256, 512};
    const double pi = 3.14159265358979323846;
    int bad = 0;
    for (int n : lens)
        for (int rows : {16, 64})
            for (int kind = 0; kind < 4; kind++)
                for (int epi = 0; epi < 2; epi++) {
                    if (epi == 1 && kind != 0) continue;
                    for (int os : {1, 2}) {
                        if (epi == 1 && os == 2) continue;
                        std::vector<float> in((size_t)rows * n), out((size_t)n * rows * os), out2(epi ? (size_t)n * rows : 1);
                        std::vector<float> tw(2 * n), dtw(2 * n), sx(2 * rows), sy(2 * n);
                        for (int k = 0; k < n; k++) {
                            tw[2 * k] = (float)std::cos(2 * pi * k / n); tw[2 * k + 1] = (float)-std::sin(2 * pi * k / n);
                            dtw[2 * k] = (float)std::cos(pi * k / (2.0 * n)); dtw[2 * k + 1] = (float)-std::sin(pi * k / (2.0 * n));
                        }
                        for (auto& v : in) v = (float)(rand() / (double)RAND_MAX);
                        for (int i = 0; i < rows; i++) { double w = pi * i / rows; sx[2 * i] = (float)w; sx[2 * i + 1] = (float)(2 * (1 - std::cos(w))); }
                        for (int i = 0; i < n; i++) { double w = pi * i / n; sy[2 * i] = (float)w; sy[2 * i + 1] = (float)(2 * (1 - std::cos(w))); }
                        DevState st{0};
                        DesignDev D{rows, n, sx.data(), sy.data()};
                        FftJob J{in.data(), out.data(), epi ? out2.data() : nullptr, tw.data(), dtw.data(), &st,
                                 n, rows, std::min(rows, fft_block_rows(n)), kind, epi, rows, os, 0};
                        std::vector<FftJob> jobs{J};
                        std::vector<int32_t> fblk;
                        for (int r = 0; r < rows; r += J.rpb) { fblk.push_back(0); fblk.push_back(r); }
                        run_pass(jobs, fblk, &D);
                        // direct transforms
                        double err = 0, mx = 0;
                        for (int r = 0; r < rows; r++)
                            for (int j = 0; j < n; j++) {
                                double y = 0;
                                const float* x = &in[(size_t)r * n];
                                for (int k = 0; k < n; k++) {
                                    double m;
                                    if (kind == 0) m = std::cos(pi * j * (2.0 * k + 1) / (2.0 * n));
                                    else if (kind == 1) m = k == 0 ? 0.5 : std::cos(pi * k * (2.0 * j + 1) / (2.0 * n));
                                    else if (kind == 2) m = k < n - 1 ? std::sin(pi * (k + 1.0) * (2.0 * j + 1) / (2.0 * n)) : ((j & 1) ? -0.5 : 0.5);
                                    else m = k == 0 ? 0.0 : -std::sin(pi * k * (2.0 * j + 1) / (2.0 * n));
                                    y += m * x[k];
                                }
                                double got;
                                if (epi == 1) {
                                    double sc = 4.0 / ((double)rows * n); if (r == 0) sc *= 0.5; if (j == 0) sc *= 0.5;
                                    double den = sx[2 * r + 1] + (double)sy[2 * j + 1];
                                    y = den < 1e-30 ? 0.0 : y * sc / den * sx[2 * r];
                                    got = out[(size_t)j * rows + r];
                                } else got = out[((size_t)j * rows + r) * os];
                                err = std::max(err, std::fabs(got - y)); mx = std::max(mx, std::fabs(y));
                            }
                        const double rel = err / std::max(mx, 1e-30);
                        if (rel > 2e-6) { bad++; printf("FAIL n=%d rows=%d kind=%d epi=%d os=%d rel=%.2e\n", n, rows, kind, epi, os, rel); }
                    }
                }
    printf(bad ? "emulation: %d failures\n" : "emulation ok (%d failures)\n", bad);
    return bad != 0;
}
