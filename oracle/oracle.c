/*
 * ElectroPlace oracle -- FP64 restatement of the reference placer.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests, the smoke check and the
 * benchmark's CPU-baseline leg as the checker; never by the product.
 *
 * Reference files restated (paths relative to crates/nextpnr/src/):
 *   placer/solver/wa.rs, placer/solver/nesterov.rs,
 *   placer/electro_place/{mod.rs,density.rs,config.rs},
 *   placer/common.rs (add_wa_wirelength_gradient, compute_pin_weights,
 *   clamp_positions, gradient_norm, NesterovLoopState),
 *   metrics/wirelength.rs, context/rng.rs.
 * The 1-D transforms follow the published rustdct 0.7.1 definitions
 * (unnormalised DCT-II, DCT-III with X0/2, DST-III with (-1)^k x_{N-1}/2);
 * rustdct itself is not vendored in the reference, so they are restated
 * here as O(N log N) FFT algorithms (Makhoul's DCT-II, its inverse, and
 * Bluestein's chirp-z for lengths that are not powers of two).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------ */
/* wa.rs                                                               */
/* ------------------------------------------------------------------ */

#define EXP_CLAMP_MIN (-3000.0) /* wa.rs:14 */

/* wa.rs:17-47 */
double orc_wa_axis_value(const double* c, int n, double a) {
    if (n < 2) return 0.0;
    double center = 0.0;
    for (int i = 0; i < n; i++) center += c[i];
    center /= (double)n;
    double smax = 0, xmax = 0, smin = 0, xmin = 0;
    for (int i = 0; i < n; i++) {
        double ma = a * (c[i] - center);
        if (ma < EXP_CLAMP_MIN) ma = EXP_CLAMP_MIN;
        double me = exp(ma);
        smax += me;
        xmax += c[i] * me;
        double na = a * (center - c[i]);
        if (na < EXP_CLAMP_MIN) na = EXP_CLAMP_MIN;
        double ne = exp(na);
        smin += ne;
        xmin += c[i] * ne;
    }
    if (smax < 1e-30 || smin < 1e-30) return 0.0;
    return xmax / smax - xmin / smin;
}

/* wa.rs:51-101 */
void orc_wa_axis_grad(const double* c, int n, double a, double* g) {
    if (n < 2) return;
    double center = 0.0;
    for (int i = 0; i < n; i++) center += c[i];
    center /= (double)n;
    double* me = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    double* ne = me + n;
    double smax = 0, smin = 0, xmax = 0, xmin = 0;
    for (int i = 0; i < n; i++) {
        double ma = a * (c[i] - center);
        if (ma < EXP_CLAMP_MIN) ma = EXP_CLAMP_MIN;
        me[i] = exp(ma);
        smax += me[i];
        xmax += c[i] * me[i];
        double na = a * (center - c[i]);
        if (na < EXP_CLAMP_MIN) na = EXP_CLAMP_MIN;
        ne[i] = exp(na);
        smin += ne[i];
        xmin += c[i] * ne[i];
    }
    if (smax < 1e-30 || smin < 1e-30) {
        free(me);
        return;
    }
    double ismax = 1.0 / smax, ismin = 1.0 / smin;
    double ismax2 = ismax * ismax, ismin2 = ismin * ismin;
    for (int i = 0; i < n; i++) {
        double x = c[i];
        double dmax = me[i] * ismax + a * me[i] * (x * smax - xmax) * ismax2;
        double dmin = ne[i] * ismin - a * ne[i] * (x * smin - xmin) * ismin2;
        g[i] += dmax - dmin;
    }
    free(me);
}

/* wa.rs:104-113 */
double orc_wa_wirelength(const double* px, const double* py, int n, double a) {
    if (n < 2) return 0.0;
    return orc_wa_axis_value(px, n, a) + orc_wa_axis_value(py, n, a);
}

static inline double pin_x(const orc_design* d, const double* x, uint32_t c) {
    return c < (uint32_t)d->n_mov ? x[c] : d->fixed_x[c - (uint32_t)d->n_mov];
}
static inline double pin_y(const orc_design* d, const double* y, uint32_t c) {
    return c < (uint32_t)d->n_mov ? y[c] : d->fixed_y[c - (uint32_t)d->n_mov];
}

/* common.rs:205-265 with net weight 1 (mod.rs:114 passes None).  The flat
 * CSR already holds each alive net's collected pins in driver-then-users
 * order, with unplaced non-movable pins dropped (common.rs:268-291). */
void orc_wa_gradient(const orc_design* d, const double* x, const double* y,
                     double a, double* gx, double* gy) {
    int maxdeg = 0;
    for (int e = 0; e < d->n_nets; e++) {
        int deg = (int)(d->net_off[e + 1] - d->net_off[e]);
        if (deg > maxdeg) maxdeg = deg;
    }
    if (maxdeg < 2) return;
    double* buf = (double*)malloc(sizeof(double) * 4 * (size_t)maxdeg);
    double *px = buf, *py = buf + maxdeg, *ggx = buf + 2 * maxdeg, *ggy = buf + 3 * maxdeg;
    for (int e = 0; e < d->n_nets; e++) {
        uint32_t b = d->net_off[e], n = d->net_off[e + 1] - b;
        if (n < 2) continue; /* common.rs:242 */
        for (uint32_t k = 0; k < n; k++) {
            uint32_t c = d->pin_cell[b + k];
            px[k] = pin_x(d, x, c);
            py[k] = pin_y(d, y, c);
            ggx[k] = 0.0;
            ggy[k] = 0.0;
        }
        orc_wa_axis_grad(px, (int)n, a, ggx);
        orc_wa_axis_grad(py, (int)n, a, ggy);
        /* common.rs:246-249: the net's weight, 1 when none is given */
        const double w = d->net_weight ? d->net_weight[e] : 1.0;
        for (uint32_t k = 0; k < n; k++) {
            uint32_t c = d->pin_cell[b + k];
            if (c < (uint32_t)d->n_mov) { /* common.rs:258-261 */
                gx[c] += w * ggx[k];
                gy[c] += w * ggy[k];
            }
        }
    }
    free(buf);
}

/* common.rs:381-407: sum of 1/degree over nets of degree >= 2. */
void orc_pin_weights(const orc_design* d, double* pinw) {
    if (d->pinw) {
        memcpy(pinw, d->pinw, sizeof(double) * (size_t)d->n_mov);
        return;
    }
    for (int i = 0; i < d->n_mov; i++) pinw[i] = 0.0;
    for (int e = 0; e < d->n_nets; e++) {
        uint32_t b = d->net_off[e], n = d->net_off[e + 1] - b;
        if (n < 2) continue;
        double w = 1.0 / (double)n;
        for (uint32_t k = 0; k < n; k++) {
            uint32_t c = d->pin_cell[b + k];
            if (c < (uint32_t)d->n_mov) pinw[c] += w;
        }
    }
}

/* ------------------------------------------------------------------ */
/* density.rs                                                          */
/* ------------------------------------------------------------------ */

/* Half side of a cell's square footprint, rounded through float exactly as
 * the product stores it.  Area 1 gives 0.5: the reference's unit square. */
static inline double half_side(float area) {
    return (double)(float)(0.5 * sqrt((double)area));
}

/* density.rs:20-37, extended: a cell of area A is a square of side sqrt(A)
 * and bin b of Bx covers tiles [b*W/Bx, (b+1)*W/Bx).  Coordinates are
 * mapped to bin units by u = (x -/+ h) * (Bx/W); with Bx == W the factor
 * is 1.0 and the expression is the reference's x_lo = cx - 0.5.
 * out = {bx_lo, bx_hi, by_lo, by_hi} (hi exclusive), box = {u_lo,u_hi,v_lo,v_hi}. */
void orc_bin_range(const orc_design* d, double cx, double cy, double area,
                   int32_t out[4], double box[4]) {
    double hs = half_side((float)area);
    double rx = (double)d->Bx / (double)d->W, ry = (double)d->By / (double)d->H;
    double xl = (cx - hs) * rx, xh = (cx + hs) * rx;
    double yl = (cy - hs) * ry, yh = (cy + hs) * ry;
    int32_t bxl = (int32_t)floor(xl);
    if (bxl < 0) bxl = 0;
    double cxh = ceil(xh);
    int32_t bxh = cxh <= 0.0 ? 0 : (cxh >= (double)d->Bx ? d->Bx : (int32_t)cxh);
    int32_t byl = (int32_t)floor(yl);
    if (byl < 0) byl = 0;
    double cyh = ceil(yh);
    int32_t byh = cyh <= 0.0 ? 0 : (cyh >= (double)d->By ? d->By : (int32_t)cyh);
    out[0] = bxl; out[1] = bxh; out[2] = byl; out[3] = byh;
    box[0] = xl; box[1] = xh; box[2] = yl; box[3] = yh;
}

/* density.rs:41-48 */
static inline double overlap_area(const double box[4], int bx, int by) {
    double ox = fmin(box[1], (double)bx + 1.0) - fmax(box[0], (double)bx);
    if (ox < 0.0) ox = 0.0;
    double oy = fmin(box[3], (double)by + 1.0) - fmax(box[2], (double)by);
    if (oy < 0.0) oy = 0.0;
    return ox * oy;
}

/* density.rs:54-74 (row-major, by * Bx + bx). */
void orc_density(const orc_design* d, const double* x, const double* y, double* dens) {
    memset(dens, 0, sizeof(double) * (size_t)d->Bx * (size_t)d->By);
    for (int i = 0; i < d->n_mov; i++) {
        int32_t r[4];
        double box[4];
        orc_bin_range(d, x[i], y[i], d->area ? d->area[i] : 1.0, r, box);
        for (int by = r[2]; by < r[3]; by++)
            for (int bx = r[0]; bx < r[1]; bx++)
                dens[(size_t)by * d->Bx + bx] += overlap_area(box, bx, by);
    }
}

/* density.rs:79-86 */
double orc_overlap(const double* dens, int64_t n) {
    double total = 0.0;
    for (int64_t i = 0; i < n; i++) total += dens[i];
    if (total < 1e-30) return 0.0;
    double of = 0.0;
    for (int64_t i = 0; i < n; i++) {
        double v = dens[i] - 1.0;
        of += v > 0.0 ? v : 0.0;
    }
    return of / total;
}

/* ---- FFT machinery for the transforms ---- */

/* A per-length plan: twiddles for the radix-2 FFT and, for lengths that are
 * not powers of two, Bluestein's chirp and the FFT of its conjugate.  Like
 * the reference's DctPlanner (density.rs:110), a plan is built once per
 * field evaluation and reused by every row and column. */
typedef struct fft_plan {
    int n, m;          /* length; padded power-of-two length (Bluestein) */
    double *tw_re, *tw_im;   /* e^{-2 pi i k / p} for the power-of-two FFT length p */
    double *ch_re, *ch_im;   /* chirp w_k = e^{-i pi k^2 / n} */
    double *bf_re, *bf_im;   /* FFT_m of conj chirp, wrapped */
    double *dc_re, *dc_im;   /* e^{-i pi k / 2n}: DCT post-twiddle */
    double *work;            /* scratch: 4*m doubles */
} fft_plan;

static void fft_pow2_tw(double* re, double* im, int n, const double* twr, const double* twi,
                        int tw_n, int inverse) {
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1, stride = tw_n / len;
        for (int i = 0; i < n; i += len) {
            for (int k = 0; k < half; k++) {
                double wr = twr[k * stride], wi = inverse ? -twi[k * stride] : twi[k * stride];
                double ur = re[i + k], ui = im[i + k];
                double vr = re[i + k + half] * wr - im[i + k + half] * wi;
                double vi = re[i + k + half] * wi + im[i + k + half] * wr;
                re[i + k] = ur + vr; im[i + k] = ui + vi;
                re[i + k + half] = ur - vr; im[i + k + half] = ui - vi;
            }
        }
    }
    if (inverse)
        for (int i = 0; i < n; i++) { re[i] /= n; im[i] /= n; }
}

static void plan_init(fft_plan* p, int n) {
    memset(p, 0, sizeof(*p));
    p->n = n;
    int pow2 = (n & (n - 1)) == 0;
    int m = n;
    if (!pow2) { m = 1; while (m < 2 * n - 1) m <<= 1; }
    p->m = m;
    p->tw_re = (double*)malloc(sizeof(double) * (size_t)m);
    p->tw_im = (double*)malloc(sizeof(double) * (size_t)m);
    for (int k = 0; k < m / 2 + 1 && k < m; k++) {
        p->tw_re[k] = cos(2.0 * M_PI * k / m);
        p->tw_im[k] = -sin(2.0 * M_PI * k / m);
    }
    p->dc_re = (double*)malloc(sizeof(double) * (size_t)n);
    p->dc_im = (double*)malloc(sizeof(double) * (size_t)n);
    for (int k = 0; k < n; k++) {
        p->dc_re[k] = cos(M_PI * k / (2.0 * n));
        p->dc_im[k] = -sin(M_PI * k / (2.0 * n));
    }
    p->work = (double*)malloc(sizeof(double) * 4 * (size_t)m);
    if (!pow2) {
        p->ch_re = (double*)malloc(sizeof(double) * (size_t)n);
        p->ch_im = (double*)malloc(sizeof(double) * (size_t)n);
        for (int k = 0; k < n; k++) {
            long long q = ((long long)k * k) % (2LL * n);
            p->ch_re[k] = cos(M_PI * (double)q / n);
            p->ch_im[k] = -sin(M_PI * (double)q / n);
        }
        p->bf_re = (double*)calloc((size_t)m, sizeof(double));
        p->bf_im = (double*)calloc((size_t)m, sizeof(double));
        p->bf_re[0] = p->ch_re[0]; p->bf_im[0] = -p->ch_im[0];
        for (int k = 1; k < n; k++) {
            p->bf_re[k] = p->bf_re[m - k] = p->ch_re[k];
            p->bf_im[k] = p->bf_im[m - k] = -p->ch_im[k];
        }
        fft_pow2_tw(p->bf_re, p->bf_im, m, p->tw_re, p->tw_im, m, 0);
    }
}

static void plan_free(fft_plan* p) {
    free(p->tw_re); free(p->tw_im); free(p->dc_re); free(p->dc_im); free(p->work);
    free(p->ch_re); free(p->ch_im); free(p->bf_re); free(p->bf_im);
}

/* In-place complex DFT of length p->n (Bluestein chirp-z when not a power
 * of two); the inverse, with its 1/n factor, is conj(DFT(conj(x)))/n. */
static void fft_plan_run(const fft_plan* p, double* re, double* im, int inverse) {
    int n = p->n, m = p->m;
    if (n <= 1) return;
    if (m == n) { fft_pow2_tw(re, im, n, p->tw_re, p->tw_im, n, inverse); return; }
    double *ar = p->work, *ai = p->work + m;
    double cj = inverse ? -1.0 : 1.0;
    for (int k = 0; k < n; k++) {
        double xr = re[k], xi = cj * im[k];
        ar[k] = xr * p->ch_re[k] - xi * p->ch_im[k];
        ai[k] = xr * p->ch_im[k] + xi * p->ch_re[k];
    }
    for (int k = n; k < m; k++) ar[k] = ai[k] = 0.0;
    fft_pow2_tw(ar, ai, m, p->tw_re, p->tw_im, m, 0);
    for (int k = 0; k < m; k++) {
        double r = ar[k] * p->bf_re[k] - ai[k] * p->bf_im[k];
        double i = ar[k] * p->bf_im[k] + ai[k] * p->bf_re[k];
        ar[k] = r; ai[k] = i;
    }
    fft_pow2_tw(ar, ai, m, p->tw_re, p->tw_im, m, 1);
    for (int k = 0; k < n; k++) {
        double r = ar[k] * p->ch_re[k] - ai[k] * p->ch_im[k];
        double i = ar[k] * p->ch_im[k] + ai[k] * p->ch_re[k];
        if (inverse) { re[k] = r / n; im[k] = -i / n; }
        else { re[k] = r; im[k] = i; }
    }
}

/* rustdct DCT-II: X_k = sum_n x_n cos(pi k (2n+1) / 2N), by Makhoul's
 * even/odd reordering and one N-point FFT. */
static void dct2_plan(const fft_plan* p, double* v) {
    int n = p->n;
    if (n <= 1) return;
    double *re = p->work + 2 * p->m, *im = re + n;
    for (int k = 0; k < (n + 1) / 2; k++) re[k] = v[2 * k];
    for (int k = 0; k < n / 2; k++) re[n - 1 - k] = v[2 * k + 1];
    for (int k = 0; k < n; k++) im[k] = 0.0;
    fft_plan_run(p, re, im, 0);
    for (int k = 0; k < n; k++) v[k] = re[k] * p->dc_re[k] - im[k] * p->dc_im[k];
}

/* rustdct DCT-III: x_n = X_0/2 + sum_{k>=1} X_k cos(pi k (2n+1) / 2N), which
 * is N/2 times the inverse of DCT-II; inverted Makhoul reordering. */
static void dct3_plan(const fft_plan* p, double* v) {
    int n = p->n;
    if (n <= 0) return;
    if (n == 1) { v[0] *= 0.5; return; }
    double *re = p->work + 2 * p->m, *im = re + n;
    for (int k = 0; k < n; k++) {
        double xr = v[k], xi = k == 0 ? 0.0 : -v[n - k];
        double c = p->dc_re[k], s = -p->dc_im[k]; /* e^{+i pi k / 2n} */
        re[k] = xr * c - xi * s;
        im[k] = xr * s + xi * c;
    }
    fft_plan_run(p, re, im, 1);
    double scale = 0.5 * (double)n;
    for (int k = 0; k < (n + 1) / 2; k++) v[2 * k] = re[k] * scale;
    for (int k = 0; k < n / 2; k++) v[2 * k + 1] = re[n - 1 - k] * scale;
}

/* rustdct DST-III: X_k = (-1)^k x_{N-1}/2 + sum_{n<N-1} x_n sin(pi (n+1)(2k+1)/2N)
 * = (-1)^k * DCT-III(reversed x)_k. */
static void dst3_plan(const fft_plan* p, double* v) {
    int n = p->n;
    if (n <= 0) return;
    for (int i = 0; i < n / 2; i++) {
        double t = v[i]; v[i] = v[n - 1 - i]; v[n - 1 - i] = t;
    }
    dct3_plan(p, v);
    for (int k = 1; k < n; k += 2) v[k] = -v[k];
}

void orc_dct2(double* v, int n) {
    if (n <= 1) return;
    fft_plan p; plan_init(&p, n); dct2_plan(&p, v); plan_free(&p);
}
void orc_dct3(double* v, int n) {
    if (n <= 0) return;
    fft_plan p; plan_init(&p, n); dct3_plan(&p, v); plan_free(&p);
}
void orc_dst3(double* v, int n) {
    if (n <= 0) return;
    fft_plan p; plan_init(&p, n); dst3_plan(&p, v); plan_free(&p);
}

void orc_dct2_naive(const double* in, double* out, int n) {
    for (int k = 0; k < n; k++) {
        double s = 0.0;
        for (int j = 0; j < n; j++) s += in[j] * cos(M_PI * k * (2.0 * j + 1.0) / (2.0 * n));
        out[k] = s;
    }
}
void orc_dct3_naive(const double* in, double* out, int n) {
    for (int j = 0; j < n; j++) {
        double s = 0.5 * in[0];
        for (int k = 1; k < n; k++) s += in[k] * cos(M_PI * k * (2.0 * j + 1.0) / (2.0 * n));
        out[j] = s;
    }
}
void orc_dst3_naive(const double* in, double* out, int n) {
    for (int k = 0; k < n; k++) {
        double s = ((k & 1) ? -0.5 : 0.5) * in[n - 1];
        for (int j = 0; j < n - 1; j++)
            s += in[j] * sin(M_PI * (j + 1.0) * (2.0 * k + 1.0) / (2.0 * n));
        out[k] = s;
    }
}

typedef void (*xform_fn)(const fft_plan*, double*);

static void rows_apply(double* a, int w, int h, const fft_plan* p, xform_fn f) {
    for (int r = 0; r < h; r++) f(p, a + (size_t)r * w);
}
static void cols_apply(double* a, int w, int h, const fft_plan* p, xform_fn f) {
    double* col = (double*)malloc(sizeof(double) * (size_t)h);
    for (int c = 0; c < w; c++) {
        for (int r = 0; r < h; r++) col[r] = a[(size_t)r * w + c];
        f(p, col);
        for (int r = 0; r < h; r++) a[(size_t)r * w + c] = col[r];
    }
    free(col);
}

/* Mode 1 only: the spectrum fed to each DST-III is shifted down by one
 * index and negated, x'[k-1] = -E[k], x'[N-1] = 0, so the sine of frequency
 * k multiplies coefficient k (the textbook IDXST) and the force repels.
 * This is the one deviation from the reference; mode 0 is the reference. */
static void shift_rows_neg(double* a, int w, int h) {
    for (int r = 0; r < h; r++) {
        double* row = a + (size_t)r * w;
        for (int k = 0; k + 1 < w; k++) row[k] = -row[k + 1];
        row[w - 1] = 0.0;
    }
}
static void shift_cols_neg(double* a, int w, int h) {
    for (int c = 0; c < w; c++) {
        for (int r = 0; r + 1 < h; r++) a[(size_t)r * w + c] = -a[(size_t)(r + 1) * w + c];
        a[(size_t)(h - 1) * w + c] = 0.0;
    }
}

void orc_field(const double* dens, int Bx, int By, double target_density,
               double* ex, double* ey) {
    orc_field_mode(dens, Bx, By, target_density, 0, ex, ey);
}

/* density.rs:96-166 */
void orc_field_mode(const double* dens, int Bx, int By, double target_density, int mode,
                    double* ex, double* ey) {
    size_t n = (size_t)Bx * By;
    double avg = 0.0;
    for (size_t i = 0; i < n; i++) avg += dens[i];
    avg /= (double)n;
    double target = target_density * (avg > 1e-30 ? avg : 1e-30);
    double* rho = (double*)malloc(sizeof(double) * n);
    for (size_t i = 0; i < n; i++) rho[i] = dens[i] - target;
    fft_plan px, py;
    plan_init(&px, Bx);
    plan_init(&py, By);
    rows_apply(rho, Bx, By, &px, dct2_plan); /* dct2_2d, density.rs:211-229 */
    cols_apply(rho, Bx, By, &py, dct2_plan);
    double scale = 4.0 / ((double)Bx * (double)By);
    for (int ky = 0; ky < By; ky++) {
        double wy = M_PI * ky / (double)By;
        double wy2 = 2.0 * (1.0 - cos(wy));
        for (int kx = 0; kx < Bx; kx++) {
            size_t idx = (size_t)ky * Bx + kx;
            double s = scale;
            if (kx == 0) s *= 0.5;
            if (ky == 0) s *= 0.5;
            double a = rho[idx] * s;
            double wx = M_PI * kx / (double)Bx;
            double wx2 = 2.0 * (1.0 - cos(wx));
            double denom = wx2 + wy2;
            if (denom < 1e-30) {
                ex[idx] = 0.0;
                ey[idx] = 0.0;
            } else {
                double phi = a / denom;
                ex[idx] = phi * wx;
                ey[idx] = phi * wy;
            }
        }
    }
    if (mode == 1) {
        shift_rows_neg(ex, Bx, By);
        shift_cols_neg(ey, Bx, By);
    }
    rows_apply(ex, Bx, By, &px, dst3_plan); /* idst_rows, density.rs:256-262 */
    cols_apply(ex, Bx, By, &py, dct3_plan); /* idct_cols, density.rs:241-253 */
    rows_apply(ey, Bx, By, &px, dct3_plan); /* idct_rows, density.rs:232-238 */
    cols_apply(ey, Bx, By, &py, dst3_plan); /* idst_cols, density.rs:265-277 */
    plan_free(&px);
    plan_free(&py);
    free(rho);
}

/* density.rs:171-206 */
void orc_density_gradient(const orc_design* d, const double* x, const double* y,
                          const double* ex, const double* ey, double* gx, double* gy) {
    for (int i = 0; i < d->n_mov; i++) {
        int32_t r[4];
        double box[4];
        orc_bin_range(d, x[i], y[i], d->area ? d->area[i] : 1.0, r, box);
        double fx = 0, fy = 0, wt = 0;
        for (int by = r[2]; by < r[3]; by++)
            for (int bx = r[0]; bx < r[1]; bx++) {
                double w = overlap_area(box, bx, by);
                if (w > 0.0) {
                    size_t idx = (size_t)by * d->Bx + bx;
                    fx += w * ex[idx];
                    fy += w * ey[idx];
                    wt += w;
                }
            }
        if (wt > 1e-10) {
            gx[i] += fx / wt;
            gy[i] += fy / wt;
        }
    }
}

/* ------------------------------------------------------------------ */
/* nesterov.rs                                                         */
/* ------------------------------------------------------------------ */

void orc_nesterov_init(orc_nesterov* s, int n, double step, double* xbuf, double* xpbuf) {
    s->n = n;
    s->x = xbuf;
    s->x_prev = xpbuf;
    s->step_size = step;
    s->iter = 0;
    s->a = 1.0;
    for (int i = 0; i < n; i++) xbuf[i] = xpbuf[i] = 0.0;
}

/* nesterov.rs:114-118 */
void orc_nesterov_set_positions(orc_nesterov* s, const double* x) {
    memcpy(s->x, x, sizeof(double) * (size_t)s->n);
    memcpy(s->x_prev, x, sizeof(double) * (size_t)s->n);
}

/* nesterov.rs:75-82 */
double orc_nesterov_momentum(const orc_nesterov* s) {
    if (s->iter == 0) return 0.0;
    double an = (1.0 + sqrt(1.0 + 4.0 * s->a * s->a)) / 2.0;
    return (s->a - 1.0) / an;
}

/* nesterov.rs:95-101 */
void orc_nesterov_look_ahead(const orc_nesterov* s, double* y) {
    double m = orc_nesterov_momentum(s);
    for (int i = 0; i < s->n; i++) y[i] = s->x[i] + m * (s->x[i] - s->x_prev[i]);
}

/* nesterov.rs:41-64 */
double orc_nesterov_step(orc_nesterov* s, const double* g) {
    s->iter += 1;
    double an = (1.0 + sqrt(1.0 + 4.0 * s->a * s->a)) / 2.0;
    double m = (s->a - 1.0) / an;
    s->a = an;
    double sq = 0.0;
    for (int i = 0; i < s->n; i++) {
        double xo = s->x[i];
        double yv = s->x[i] + m * (s->x[i] - s->x_prev[i]);
        double xn = yv - s->step_size * g[i];
        s->x_prev[i] = xo;
        s->x[i] = xn;
        double dl = xn - xo;
        sq += dl * dl;
    }
    return sqrt(sq);
}

static inline double clampd(double v, double lo, double hi) {
    return v < lo ? lo : (v > hi ? hi : v);
}

/* nesterov.rs:67-72 */
void orc_nesterov_clamp(orc_nesterov* s, double lo, double hi) {
    for (int i = 0; i < s->n; i++) {
        s->x[i] = clampd(s->x[i], lo, hi);
        s->x_prev[i] = clampd(s->x_prev[i], lo, hi);
    }
}

/* nesterov.rs:199-218; returns 0 for None. */
int orc_nesterov_bb(const orc_nesterov* s, const double* gp, const double* g, double* out) {
    double dxdx = 0.0, dxdg = 0.0;
    for (int i = 0; i < s->n; i++) {
        double dx = s->x[i] - s->x_prev[i];
        double dg = g[i] - gp[i];
        dxdx += dx * dx;
        dxdg += dx * dg;
    }
    if (fabs(dxdg) < 1e-30) return 0;
    *out = fabs(dxdx / dxdg);
    return 1;
}

/* nesterov.rs:171-191 */
double orc_nesterov_lipschitz(const orc_nesterov* s, const double* gp, const double* g) {
    double a = 0.0, b = 0.0;
    for (int i = 0; i < s->n; i++) {
        double dx = s->x[i] - s->x_prev[i];
        double dg = g[i] - gp[i];
        a += dx * dx;
        b += dg * dg;
    }
    double gn = sqrt(b);
    return gn > 1e-30 ? sqrt(a) / gn : s->step_size;
}

/* nesterov.rs:125-139 */
void orc_nesterov_adaptive_restart(orc_nesterov* s, const double* g) {
    double dot = 0.0;
    for (int i = 0; i < s->n; i++) dot += g[i] * (s->x[i] - s->x_prev[i]);
    if (dot > 0.0) {
        memcpy(s->x_prev, s->x, sizeof(double) * (size_t)s->n);
        s->iter = 0;
        s->a = 1.0;
    }
}

/* ------------------------------------------------------------------ */
/* metrics                                                             */
/* ------------------------------------------------------------------ */

/* Continuous HPWL: wirelength.rs:10-43 evaluated on continuous positions
 * (every pin of a net with >= 2 pins; fixed pins at their locations). */
double orc_hpwl(const orc_design* d, const double* x, const double* y) {
    double total = 0.0;
    for (int e = 0; e < d->n_nets; e++) {
        uint32_t b = d->net_off[e], n = d->net_off[e + 1] - b;
        if (n < 2) continue;
        double xmn = 1e300, xmx = -1e300, ymn = 1e300, ymx = -1e300;
        for (uint32_t k = 0; k < n; k++) {
            uint32_t c = d->pin_cell[b + k];
            double px = pin_x(d, x, c), py = pin_y(d, y, c);
            xmn = fmin(xmn, px); xmx = fmax(xmx, px);
            ymn = fmin(ymn, py); ymx = fmax(ymx, py);
        }
        total += (xmx - xmn) + (ymx - ymn);
    }
    return total;
}

/* wirelength.rs:10-51 on integer (legalized) locations: driver = first pin,
 * nets without users contribute 0. */
double orc_total_hpwl_int(const orc_design* d, const int32_t* cix, const int32_t* ciy) {
    double total = 0.0;
    for (int e = 0; e < d->n_nets; e++) {
        uint32_t b = d->net_off[e], n = d->net_off[e + 1] - b;
        if (n < 2) continue;
        int32_t xmn = 0x7fffffff, xmx = -0x7fffffff - 1, ymn = 0x7fffffff, ymx = -0x7fffffff - 1;
        for (uint32_t k = 0; k < n; k++) {
            uint32_t c = d->pin_cell[b + k];
            int32_t px, py;
            if (c < (uint32_t)d->n_mov) { px = cix[c]; py = ciy[c]; }
            else { px = (int32_t)d->fixed_x[c - d->n_mov]; py = (int32_t)d->fixed_y[c - d->n_mov]; }
            if (px < xmn) xmn = px;
            if (px > xmx) xmx = px;
            if (py < ymn) ymn = py;
            if (py > ymx) ymx = py;
        }
        total += (double)((xmx - xmn) + (ymx - ymn));
    }
    return total;
}

/* mod.rs:206-256: unbind all movable cells, then for each cell in solver
 * index order pick the available BEL of its bucket with the smallest
 * squared distance (strict <, so the first minimum in bucket order wins).
 * Bucket order is BEL index order (buckets.rs:37-43).  bel_taken holds the
 * BELs bound by fixed cells on entry and is updated.  Returns 0 or -2
 * (NoBelsAvailable). */
int orc_legalize(int n_cells, const double* x, const double* y, const int32_t* cell_bucket,
                 int n_bels, const int32_t* bel_x, const int32_t* bel_y,
                 const int32_t* bel_bucket, uint8_t* bel_taken, int32_t* cell_bel,
                 double* displacement) {
    double disp = 0.0;
    for (int i = 0; i < n_cells; i++) {
        int best = -1;
        double bc = INFINITY;
        for (int b = 0; b < n_bels; b++) {
            if (bel_bucket[b] != cell_bucket[i] || bel_taken[b]) continue;
            double dx = (double)bel_x[b] - x[i];
            double dy = (double)bel_y[b] - y[i];
            double c = dx * dx + dy * dy;
            if (c < bc) { bc = c; best = b; }
        }
        if (best < 0) return -2;
        bel_taken[best] = 1;
        cell_bel[i] = best;
        disp += bc;
    }
    if (displacement) *displacement = disp;
    return 0;
}

/* ------------------------------------------------------------------ */
/* rng.rs                                                              */
/* ------------------------------------------------------------------ */

/* rng.rs:26-31 */
uint64_t orc_rng_next(uint64_t* st) {
    uint64_t s = *st;
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    *st = s;
    return s;
}

/* rng.rs:52-61 (next_range = (u64 as u32) % max, rng.rs:45-47) */
void orc_rng_shuffle_u32(uint64_t* st, uint32_t* v, int n) {
    if (n <= 1) return;
    for (int i = n - 1; i >= 1; i--) {
        uint32_t j = ((uint32_t)orc_rng_next(st)) % (uint32_t)(i + 1);
        uint32_t t = v[i]; v[i] = v[j]; v[j] = t;
    }
}

/* ------------------------------------------------------------------ */
/* mod.rs:65-204, the continuous loop                                  */
/* ------------------------------------------------------------------ */

static double grad_norm(const double* gx, const double* gy, int n) {
    double s = 0.0; /* common.rs:418-425: x then y chained */
    for (int i = 0; i < n; i++) s += gx[i] * gx[i];
    for (int i = 0; i < n; i++) s += gy[i] * gy[i];
    return sqrt(s);
}

/* Called on the legalization iterations of mod.rs:176 (and on the one a
 * per-iteration stop rule ends at) with the solver positions after the step;
 * a nonzero return aborts the loop with that status. */
typedef int (*orc_check_fn)(void* user, int iter, const double* x, const double* y);

static int place_core(const orc_design* d, const orc_cfg* cfg, double* x_out, double* y_out,
                      orc_iter_rec* log, int log_cap, int* n_log, orc_result* res,
                      orc_check_fn on_check, void* user) {
    int n = d->n_mov;
    if (n_log) *n_log = 0;
    memset(res, 0, sizeof(*res));
    if (n == 0) return 0; /* mod.rs:80-82 */
    double max_x = (double)(d->W - 1), max_y = (double)(d->H - 1); /* mod.rs:73-74 */
    size_t nb = (size_t)d->Bx * d->By;
    double* mem = (double*)calloc((size_t)14 * n + 3 * nb, sizeof(double));
    double *cx = mem, *cy = mem + n, *gx = mem + 2 * n, *gy = mem + 3 * n;
    double *dgx = mem + 4 * n, *dgy = mem + 5 * n, *pgx = mem + 6 * n, *pgy = mem + 7 * n;
    double *pinw = mem + 8 * n, *sx = mem + 9 * n, *sxp = mem + 10 * n, *sy = mem + 11 * n;
    double *syp = mem + 12 * n;
    double *dens = mem + 14 * (size_t)n, *ex = dens + nb, *ey = ex + nb;
    orc_nesterov nx, ny;
    orc_nesterov_init(&nx, n, cfg->nesterov_step_size, sx, sxp);
    orc_nesterov_init(&ny, n, cfg->nesterov_step_size, sy, syp);
    orc_nesterov_set_positions(&nx, d->x0); /* mod.rs:86-91 */
    orc_nesterov_set_positions(&ny, d->y0);
    orc_pin_weights(d, pinw); /* mod.rs:98 */
    double lambda = 0.0;
    int lambda_init = 0;
    int nlog = 0, converged = 0, it = 0;
    double last_ov = 0.0;
    for (it = 0; it < cfg->max_iters; it++) {
        orc_nesterov_look_ahead(&nx, cx); /* mod.rs:105-107 */
        orc_nesterov_look_ahead(&ny, cy);
        for (int i = 0; i < n; i++) {
            cx[i] = clampd(cx[i], 0.0, max_x);
            cy[i] = clampd(cy[i], 0.0, max_y);
        }
        memset(gx, 0, sizeof(double) * n);
        memset(gy, 0, sizeof(double) * n);
        orc_wa_gradient(d, cx, cy, cfg->wl_coeff, gx, gy); /* mod.rs:110-115 */
        orc_density(d, cx, cy, dens);                       /* mod.rs:118 */
        double ov = orc_overlap(dens, (int64_t)nb);         /* mod.rs:119 */
        last_ov = ov;
        orc_field_mode(dens, d->Bx, d->By, cfg->target_density, cfg->field_mode, ex, ey); /* mod.rs:121-123 */
        memset(dgx, 0, sizeof(double) * n);
        memset(dgy, 0, sizeof(double) * n);
        orc_density_gradient(d, cx, cy, ex, ey, dgx, dgy); /* mod.rs:126-129 */
        double wln = 0.0, dnn = 0.0;
        if (!lambda_init) { /* mod.rs:131-143 */
            wln = grad_norm(gx, gy, n);
            dnn = grad_norm(dgx, dgy, n);
            if (dnn > 1e-30) {
                lambda = 0.1 * wln / dnn;
                lambda_init = 1;
            }
        } else if (lambda < 50.0) {
            lambda *= 1.025;
        } else {
            lambda += 1.0;
        }
        for (int i = 0; i < n; i++) { /* mod.rs:146-149 */
            gx[i] += lambda * dgx[i];
            gy[i] += lambda * dgy[i];
        }
        for (int i = 0; i < n; i++) { /* mod.rs:152-156 */
            double pc = pinw[i] + lambda;
            if (pc < 1.0) pc = 1.0;
            gx[i] /= pc;
            gy[i] /= pc;
        }
        if (it > 0) { /* mod.rs:159-166 */
            double bb;
            if (orc_nesterov_bb(&nx, pgx, gx, &bb)) nx.step_size = clampd(bb, 1e-4, 1.0);
            if (orc_nesterov_bb(&ny, pgy, gy, &bb)) ny.step_size = clampd(bb, 1e-4, 1.0);
        }
        memcpy(pgx, gx, sizeof(double) * n); /* mod.rs:167 */
        memcpy(pgy, gy, sizeof(double) * n);
        double stx = orc_nesterov_step(&nx, gx); /* mod.rs:169-173 */
        double sty = orc_nesterov_step(&ny, gy);
        orc_nesterov_clamp(&nx, 0.0, max_x);
        orc_nesterov_clamp(&ny, 0.0, max_y);
        int check = cfg->stop_every_iter ? 1
                    : (it % cfg->legalize_interval == 0 || it == cfg->max_iters - 1);
        double hp = 0.0;
        if (check) hp = orc_hpwl(d, nx.x, ny.x);
        if (log && nlog < log_cap) {
            orc_iter_rec* r = &log[nlog++];
            r->iter = it;
            r->checked = check;
            r->overflow = ov;
            r->lambda = lambda;
            r->step_size_x = nx.step_size;
            r->step_size_y = ny.step_size;
            r->step_norm_x = stx;
            r->step_norm_y = sty;
            r->hpwl = hp;
            r->wl_norm = wln;
            r->den_norm = dnn;
        }
        int stop = check && it >= cfg->min_iters && ov < cfg->overflow_target;
        int leg = it % cfg->legalize_interval == 0 || it == cfg->max_iters - 1;
        if (on_check && (leg || stop)) { /* mod.rs:176-183: legalize, HPWL, record_metric */
            int rc = on_check(user, it, nx.x, ny.x);
            if (rc) {
                free(mem);
                return rc;
            }
        }
        if (stop) { /* mod.rs:192-195 */
            converged = 1;
            it++;
            break;
        }
    }
    if (x_out) memcpy(x_out, nx.x, sizeof(double) * n);
    if (y_out) memcpy(y_out, ny.x, sizeof(double) * n);
    res->iters = it;
    res->converged = converged;
    res->overflow = last_ov;
    res->hpwl = orc_hpwl(d, nx.x, ny.x);
    res->lambda = lambda;
    if (n_log) *n_log = nlog;
    free(mem);
    return 0;
}

int orc_place(const orc_design* d, const orc_cfg* cfg, double* x_out, double* y_out,
              orc_iter_rec* log, int log_cap, int* n_log, orc_result* res) {
    return place_core(d, cfg, x_out, y_out, log, log_cap, n_log, res, NULL, NULL);
}

/* ------------------------------------------------------------------ */
/* legalization with cluster children and the best snapshot            */
/* ------------------------------------------------------------------ */

/* mod.rs:206-256 with place_cluster_children (common.rs:319-374): every
 * movable cell in solver order takes the first minimum of the squared
 * distance over the free BELs of its bucket, scanned in bucket (BEL index)
 * order; right after it, each of its cluster children binds the first free
 * BEL of the child's bucket at root location + (dx, dy), else the first free
 * BEL of that bucket.  taken holds the BELs bound on entry and is updated.
 * Returns 0, -2 (NoBelsAvailable) or -3 (PlacementFailed). */
int orc_legalize_full(const orc_design* d, const orc_legal* lp, const double* x, const double* y,
                      uint8_t* taken, int32_t* cell_bel, int32_t* child_bel, double* displacement) {
    double disp = 0.0;
    int ch = 0;
    for (int i = 0; i < d->n_mov; i++) {
        int best = -1;
        double bc = INFINITY;
        for (int b = 0; b < lp->n_bels; b++) {
            if (lp->bel_bucket[b] != lp->cell_bucket[i] || taken[b]) continue;
            double dx = (double)lp->bel_x[b] - x[i];
            double dy = (double)lp->bel_y[b] - y[i];
            double c = dx * dx + dy * dy;
            if (c < bc) { bc = c; best = b; }
        }
        if (best < 0) return -2;
        taken[best] = 1;
        cell_bel[i] = best;
        disp += bc;
        for (; ch < lp->n_children && lp->child_root[ch] == i; ch++) {
            int tx = lp->bel_x[best] + lp->child_dx[ch], ty = lp->bel_y[best] + lp->child_dy[ch];
            int cb = -1;
            for (int b = 0; b < lp->n_bels && cb < 0; b++)
                if (lp->bel_bucket[b] == lp->child_bucket[ch] && !taken[b] && lp->bel_x[b] == tx &&
                    lp->bel_y[b] == ty)
                    cb = b;
            for (int b = 0; b < lp->n_bels && cb < 0; b++)
                if (lp->bel_bucket[b] == lp->child_bucket[ch] && !taken[b]) cb = b;
            if (cb < 0) return -3;
            taken[cb] = 1;
            child_bel[ch] = cb;
        }
    }
    if (displacement) *displacement = disp;
    return 0;
}

/* total_hpwl (wirelength.rs:10-51) of a binding: per net with a driver and
 * a user, (max - min) x + y over its placed pins; movable cells at their
 * BELs, cluster children at theirs, other fixed cells at their locations. */
double orc_legal_hpwl(const orc_design* d, const orc_legal* lp, const int32_t* cell_bel,
                      const int32_t* child_bel) {
    int32_t* fx = (int32_t*)malloc(sizeof(int32_t) * (size_t)(d->n_fixed + 1));
    int32_t* fy = (int32_t*)malloc(sizeof(int32_t) * (size_t)(d->n_fixed + 1));
    for (int j = 0; j < d->n_fixed; j++) {
        fx[j] = (int32_t)d->fixed_x[j];
        fy[j] = (int32_t)d->fixed_y[j];
    }
    for (int k = 0; k < lp->n_children; k++)
        if (lp->child_fixed[k] >= 0) {
            fx[lp->child_fixed[k]] = lp->bel_x[child_bel[k]];
            fy[lp->child_fixed[k]] = lp->bel_y[child_bel[k]];
        }
    double total = 0.0;
    for (int e = 0; e < d->n_nets; e++) {
        uint32_t b = d->net_off[e], n = d->net_off[e + 1] - b;
        if (n < 2) continue;
        if (lp->net_has_driver && !lp->net_has_driver[e]) continue;
        int32_t xmn = 0x7fffffff, xmx = -0x7fffffff - 1, ymn = 0x7fffffff, ymx = -0x7fffffff - 1;
        for (uint32_t k = 0; k < n; k++) {
            uint32_t c = d->pin_cell[b + k];
            int32_t px, py;
            if (c < (uint32_t)d->n_mov) { px = lp->bel_x[cell_bel[c]]; py = lp->bel_y[cell_bel[c]]; }
            else { px = fx[c - d->n_mov]; py = fy[c - d->n_mov]; }
            if (px < xmn) xmn = px;
            if (px > xmx) xmx = px;
            if (py < ymn) ymn = py;
            if (py > ymx) ymx = py;
        }
        total += (double)((xmx - xmn) + (ymx - ymn));
    }
    free(fx);
    free(fy);
    return total;
}

typedef struct {
    const orc_design* d;       /* fixed_x / fixed_y point at fx / fy */
    const orc_legal* lp;
    double *fx, *fy;           /* the fixed locations the loop reads (children move) */
    uint8_t* taken;
    int32_t *cb, *kb;
    double *bx, *by;           /* best positions */
    double best;
    int best_iter, count;
} leg_state;

static int leg_check(void* user, int it, const double* x, const double* y) {
    leg_state* S = (leg_state*)user;
    const orc_design* d = S->d;
    const orc_legal* lp = S->lp;
    /* mod.rs:177-179: positions of the solvers, clamped (already in range) */
    for (int b = 0; b < lp->n_bels; b++) S->taken[b] = lp->bel_taken ? lp->bel_taken[b] : 0;
    int rc = orc_legalize_full(d, lp, x, y, S->taken, S->cb, S->kb, NULL);
    if (rc) return rc;
    S->count++;
    /* the children now sit on their new BELs, for the HPWL and the next iterations */
    for (int k = 0; k < lp->n_children; k++)
        if (lp->child_fixed[k] >= 0) {
            S->fx[lp->child_fixed[k]] = lp->bel_x[S->kb[k]];
            S->fy[lp->child_fixed[k]] = lp->bel_y[S->kb[k]];
        }
    double h = orc_legal_hpwl(d, lp, S->cb, S->kb); /* wirelength.rs:46-51 */
    if (h < S->best) { /* common.rs:485-495 */
        S->best = h;
        S->best_iter = it;
        memcpy(S->bx, x, sizeof(double) * d->n_mov);
        memcpy(S->by, y, sizeof(double) * d->n_mov);
    }
    return 0;
}

/* place_electro (mod.rs:65-204) from the supplied initial positions: the
 * continuous loop, at each legalization iteration the legalization, its
 * HPWL and the best snapshot, then the final binding legalize(best)
 * (mod.rs:199).  Returns 0 or the legalization's status. */
int orc_place_legalized(const orc_design* d0, const orc_cfg* cfg, const orc_legal* lp,
                        int32_t* cell_bel, int32_t* child_bel, double* best_hpwl, int* best_iter,
                        int* legalizations, orc_result* res) {
    int n = d0->n_mov;
    orc_design d = *d0;
    leg_state S;
    memset(&S, 0, sizeof(S));
    S.fx = (double*)malloc(sizeof(double) * (size_t)(d0->n_fixed + 1));
    S.fy = (double*)malloc(sizeof(double) * (size_t)(d0->n_fixed + 1));
    for (int j = 0; j < d0->n_fixed; j++) {
        S.fx[j] = d0->fixed_x[j];
        S.fy[j] = d0->fixed_y[j];
    }
    d.fixed_x = S.fx;
    d.fixed_y = S.fy;
    S.d = &d;
    S.lp = lp;
    S.taken = (uint8_t*)malloc((size_t)lp->n_bels + 1);
    S.cb = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n + 1));
    S.kb = (int32_t*)malloc(sizeof(int32_t) * (size_t)(lp->n_children + 1));
    S.bx = (double*)malloc(sizeof(double) * (size_t)(n + 1));
    S.by = (double*)malloc(sizeof(double) * (size_t)(n + 1));
    memcpy(S.bx, d0->x0, sizeof(double) * n); /* NesterovLoopState::new (common.rs:459-468) */
    memcpy(S.by, d0->y0, sizeof(double) * n);
    S.best = INFINITY;
    S.best_iter = -1;
    int rc = place_core(&d, cfg, NULL, NULL, NULL, 0, NULL, res, leg_check, &S);
    if (!rc) {
        for (int b = 0; b < lp->n_bels; b++) S.taken[b] = lp->bel_taken ? lp->bel_taken[b] : 0;
        rc = orc_legalize_full(&d, lp, S.bx, S.by, S.taken, cell_bel, child_bel, NULL);
    }
    *best_hpwl = S.best;
    *best_iter = S.best_iter;
    *legalizations = S.count;
    free(S.fx); free(S.fy); free(S.taken); free(S.cb); free(S.kb); free(S.bx); free(S.by);
    return rc;
}
