/*
 * glsu_oracle.c -- FP64 CPU oracle (test infrastructure, not product code).
 *
 * Restates the reference's FP64 CPU path for the per-frame G-LS3U chain:
 *   DPRA least squares      DPRA/dpra_cpu.cpp:130-232
 *   WFF2                    WFT/WFT2_CPU.cpp:190-390 (init :697-845, thr :977-1022)
 *   phase / ref update      DPRA/dpra_cpu.cpp:242-258
 *   frame loop              DPRA/dpra_cpu.cpp:58-89
 *   OPT padding table       WFT/WFT.h:18-32, WFT/WFT.cpp:10-23, MatLab/FFT_OptimalSize.m:1-23
 *   AIA (initial phase)     AIA/aia_cpuf.cpp:16-260, AIA/aia_cudaf.cu:290-461, MatLab/fp_aia.m
 *
 * The reference binds two third-party routines on this path (Intel MKL, version
 * not pinned in the repository: DPRA/DPRA.vcxproj <UseIntelMKL>Parallel):
 *   - the FFTW3 API (fftw_plan_dft_2d, FFTW_ESTIMATE), WFT/WFT2_CPU.cpp:745,841-843;
 *   - LAPACKE_dsysv (Bunch-Kaufman), DPRA/dpra_cpu.cpp:219.
 * They are replaced by the oracle's own mixed-radix FP64 Stockham FFT (published
 * algorithm: Stockham autosort, radix 2/3/4/5/7/generic) and an FP64 Gaussian
 * elimination with partial pivoting on the symmetric 3x3 system.  Neither is a
 * bit-level copy of MKL; both agree with it to FP64 rounding on non-singular
 * inputs.
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *   FFT            numpy.fft (pocketfft) on 2^a3^b5^c7^d and prime sizes; golden
 *                  vectors in tests/golden/fft_vectors.npz.
 *   WFF2           brute-force direct double convolution (numpy, written from
 *                  the formula F = c * sum_p w_p * T(w_p * f)) on tiny inputs;
 *                  padding independence; thr=+inf => 0; conjugation symmetry;
 *                  single-tone known answer; separable Fg == literal FFT2(gwave).
 *   threshold      closed form 6*sqrt(mean|f|^2/3) and first-call caching.
 *   OPT table      regenerated from the MATLAB rule; known answers.
 *   pair count     known answers for double and float-literal inputs.
 *   DPRA LS        noise-free closed form; numpy.linalg.solve brute force incl.
 *                  border pixels (A00 = 9 quirk); offset/gain invariance.
 *   update         closed-form cumulative sums for a WFF-free noise-free chain.
 *   AIA            the reference's Test/Phi_CPU.csv (per-pixel phase-step solutions on
 *                  Test/{1,2,3}.bmp); noise-free closed forms of each half-step and of
 *                  the loop; three-frame exact fit; shift-offset gauge; loop rule and
 *                  error metrics against half-steps driven from Python.
 * Parity unpinned: MKL's exact rounding (FFTW_ESTIMATE plans, dsysv pivots) and
 * what dsysv leaves in b on singular pixels; the oracle defines that case as
 * Delta = 0 (|det| <= 1e-10 * A00*A11*A22).
 */
#include "glsu_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ========================================================================== */
/* OPT_FFT_SIZE: even 2^a 3^b 5^c 7^d <= 4096 (MatLab/FFT_OptimalSize.m:4-16) */
/* ========================================================================== */
static int g_opt[256];
static int g_opt_n = 0;

static int cmp_int(const void *a, const void *b) {
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

static void opt_build(void) {
    if (g_opt_n) return;
    int n = 0;
    for (int i = 0; i <= 12; i++)
        for (int j = 0; j <= 12; j++)
            for (int k = 0; k <= 12; k++)
                for (int l = 0; l <= 12; l++) {
                    double t = pow(2, i) * pow(3, j) * pow(5, k) * pow(7, l);
                    if (t <= 4096 && ((long)t) % 2 == 0) g_opt[n++] = (int)t;
                }
    qsort(g_opt, n, sizeof(int), cmp_int);
    g_opt_n = n;
}

int orc_opt_table(int *out, int cap) {
    opt_build();
    for (int i = 0; i < g_opt_n && i < cap; i++) out[i] = g_opt[i];
    return g_opt_n;
}

/* lower-bound binary search, WFT/WFT.cpp:10-23 (returns -1 past the end) */
int orc_first_greater(int x) {
    opt_build();
    int first = 0, last = g_opt_n - 1;
    while (first <= last) {
        int mid = (first + last) / 2;
        if (g_opt[mid] >= x) last = mid - 1;
        else first = mid + 1;
    }
    return last + 1 == g_opt_n ? -1 : last + 1;
}

int orc_opt_size(int x) {
    int i = orc_first_greater(x);
    return i < 0 ? -1 : g_opt[i];
}

/* WFT/WFT2_CPU.cpp:222-223: int((wh - wl) * (1 / wi)) + 1, in double */
int orc_pair_count(double wl, double wh, double wi) {
    return (int)((wh - wl) * (1 / wi)) + 1;
}

/* ========================================================================== */
/* FFT: Stockham autosort, radices 4/2/3/5/7/generic, FP64                    */
/* ========================================================================== */
typedef struct {
    int n, nf;
    int f[40];
    double *tw;   /* tw[2k], tw[2k+1] = cos, sin of 2*pi*k/n */
} fftplan;

static void plan_init(fftplan *p, int n) {
    p->n = n;
    p->nf = 0;
    int m = n;
    while (m % 4 == 0) { p->f[p->nf++] = 4; m /= 4; }
    while (m % 2 == 0) { p->f[p->nf++] = 2; m /= 2; }
    for (int r = 3; m > 1; r += 2)
        while (m % r == 0) { p->f[p->nf++] = r; m /= r; }
    p->tw = (double *)malloc(sizeof(double) * 2 * (size_t)(n > 0 ? n : 1));
    for (int k = 0; k < n; k++) {
        double a = 2.0 * M_PI * (double)k / (double)n;
        p->tw[2 * k] = cos(a);
        p->tw[2 * k + 1] = sin(a);
    }
}

static void plan_free(fftplan *p) { free(p->tw); p->tw = NULL; }

/* x, work: interleaved complex of length n; sign -1 forward, +1 backward */
static void fft_exec(const fftplan *p, double *x, double *work, int sign) {
    const int n = p->n;
    if (n <= 1) return;
    double *src = x, *dst = work;
    int Ls = 1;
    int rmax = 1;
    for (int s = 0; s < p->nf; s++) if (p->f[s] > rmax) rmax = p->f[s];
    double *ubuf = (double *)malloc(sizeof(double) * 4 * (size_t)rmax);
    double *ur = ubuf, *ui = ubuf + rmax, *vr = ubuf + 2 * rmax, *vi = ubuf + 3 * rmax;
    for (int s = 0; s < p->nf; s++) {
        const int r = p->f[s];
        const int m = n / r;
        const int tstride = n / (Ls * r);
        const int rstride = n / r;
        for (int b = 0; b < m; b++) {
            const int j = b % Ls;
            for (int q = 0; q < r; q++) {
                double a = src[2 * (b + q * m)], c = src[2 * (b + q * m) + 1];
                int ti = (int)(((long)j * q * tstride) % n);
                double wr = p->tw[2 * ti], wi = sign * p->tw[2 * ti + 1];
                ur[q] = a * wr - c * wi;
                ui[q] = a * wi + c * wr;
            }
            if (r == 2) {
                vr[0] = ur[0] + ur[1]; vi[0] = ui[0] + ui[1];
                vr[1] = ur[0] - ur[1]; vi[1] = ui[0] - ui[1];
            } else if (r == 4) {
                double t0r = ur[0] + ur[2], t0i = ui[0] + ui[2];
                double t1r = ur[0] - ur[2], t1i = ui[0] - ui[2];
                double t2r = ur[1] + ur[3], t2i = ui[1] + ui[3];
                double t3r = ur[1] - ur[3], t3i = ui[1] - ui[3];
                /* multiply t3 by sign*i */
                double t3mr = -sign * t3i, t3mi = sign * t3r;
                vr[0] = t0r + t2r;  vi[0] = t0i + t2i;
                vr[2] = t0r - t2r;  vi[2] = t0i - t2i;
                vr[1] = t1r + t3mr; vi[1] = t1i + t3mi;
                vr[3] = t1r - t3mr; vi[3] = t1i - t3mi;
            } else {
                for (int k = 0; k < r; k++) {
                    double sr = 0, si = 0;
                    for (int q = 0; q < r; q++) {
                        int ti = (int)(((long)q * k % r) * rstride);
                        double wr = p->tw[2 * ti], wi = sign * p->tw[2 * ti + 1];
                        sr += ur[q] * wr - ui[q] * wi;
                        si += ur[q] * wi + ui[q] * wr;
                    }
                    vr[k] = sr; vi[k] = si;
                }
            }
            const int base = (b - j) * r + j;
            for (int q = 0; q < r; q++) {
                dst[2 * (base + q * Ls)] = vr[q];
                dst[2 * (base + q * Ls) + 1] = vi[q];
            }
        }
        double *t = src; src = dst; dst = t;
        Ls *= r;
    }
    free(ubuf);
    if (src != x) memcpy(x, src, sizeof(double) * 2 * (size_t)n);
}

int orc_fft1d(int n, double *data, int sign) {
    if (n < 1) return -1;
    fftplan p;
    plan_init(&p, n);
    double *w = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    fft_exec(&p, data, w, sign);
    free(w);
    plan_free(&p);
    return 0;
}

typedef struct {
    int rows, cols;
    fftplan pr, pc;     /* along a row (length cols), along a column (rows) */
    double *tmp;        /* rows*cols complex, transpose buffer */
    double *work;       /* max(rows, cols) complex */
} fft2ctx;

static void fft2_init(fft2ctx *c, int rows, int cols) {
    c->rows = rows;
    c->cols = cols;
    plan_init(&c->pr, cols);
    plan_init(&c->pc, rows);
    c->tmp = (double *)malloc(sizeof(double) * 2 * (size_t)rows * cols);
    c->work = (double *)malloc(sizeof(double) * 2 * (size_t)(rows > cols ? rows : cols));
}

static void fft2_free(fft2ctx *c) {
    plan_free(&c->pr);
    plan_free(&c->pc);
    free(c->tmp);
    free(c->work);
}

static void transpose_c(const double *in, double *out, int rows, int cols) {
    const int B = 32;
    for (int i0 = 0; i0 < rows; i0 += B)
        for (int j0 = 0; j0 < cols; j0 += B) {
            int i1 = i0 + B < rows ? i0 + B : rows;
            int j1 = j0 + B < cols ? j0 + B : cols;
            for (int i = i0; i < i1; i++)
                for (int j = j0; j < j1; j++) {
                    out[2 * ((size_t)j * rows + i)] = in[2 * ((size_t)i * cols + j)];
                    out[2 * ((size_t)j * rows + i) + 1] = in[2 * ((size_t)i * cols + j) + 1];
                }
        }
}

/* 2-D DFT of a rows x cols row-major array (fftw_plan_dft_2d semantics). */
static void fft2_exec(fft2ctx *c, double *x, int sign) {
    for (int i = 0; i < c->rows; i++) fft_exec(&c->pr, x + 2 * (size_t)i * c->cols, c->work, sign);
    transpose_c(x, c->tmp, c->rows, c->cols);
    for (int j = 0; j < c->cols; j++) fft_exec(&c->pc, c->tmp + 2 * (size_t)j * c->rows, c->work, sign);
    transpose_c(c->tmp, x, c->cols, c->rows);
}

int orc_fft2d(int rows, int cols, double *data, int sign) {
    if (rows < 1 || cols < 1) return -1;
    fft2ctx c;
    fft2_init(&c, rows, cols);
    fft2_exec(&c, data, sign);
    fft2_free(&c);
    return 0;
}

/* ========================================================================== */
/* Synthetic fringe generator                                                 */
/* ========================================================================== */
/* SplitMix64 finaliser (Steele, Lea, Flood 2014), chained over (seed, t, p). */
static uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

uint64_t orc_hash3(uint64_t seed, uint64_t t, uint64_t p) {
    return mix64(mix64(mix64(seed) ^ t) ^ p);
}

/* one N(0,1) sample per (seed, frame, pixel): FP64 Box-Muller, cosine branch */
double orc_normal(uint64_t seed, uint64_t t, uint64_t p) {
    uint64_t h1 = orc_hash3(seed, t, 2 * p);
    uint64_t h2 = orc_hash3(seed, t, 2 * p + 1);
    double u1 = ((double)(h1 >> 11) + 0.5) * 0x1.0p-53;
    double u2 = (double)(h2 >> 11) * 0x1.0p-53;
    return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

/* MATLAB peaks(X, Y) */
double orc_peaks(double X, double Y) {
    return 3.0 * (1.0 - X) * (1.0 - X) * exp(-X * X - (Y + 1.0) * (Y + 1.0))
         - 10.0 * (X / 5.0 - X * X * X - Y * Y * Y * Y * Y) * exp(-X * X - Y * Y)
         - exp(-(X + 1.0) * (X + 1.0) - Y * Y) / 3.0;
}

static double lin_m3p3(int k, int n) { return n > 1 ? -3.0 + 6.0 * (double)k / (double)(n - 1) : 3.0; }

/* phi(x,y) = amp * peaks(X, Y) + 2*pi*carrier*x, X along columns, Y along rows */
void orc_gen_phase(int W, int H, double amp, double carrier, double *phi) {
    for (int i = 0; i < H; i++) {
        double Y = lin_m3p3(i, H);
        for (int j = 0; j < W; j++) {
            double X = lin_m3p3(j, W);
            phi[(size_t)i * W + j] = amp * orc_peaks(X, Y) + 2.0 * M_PI * carrier * (double)j;
        }
    }
}

void orc_gen_frame(int W, int H, int t, uint64_t seed, double amp0, double amp1,
                   double carrier, double noise, double *I, double *phi) {
    double amp = amp0 + amp1 * (double)t;
    for (int i = 0; i < H; i++) {
        double Y = lin_m3p3(i, H);
        for (int j = 0; j < W; j++) {
            double X = lin_m3p3(j, W);
            size_t p = (size_t)i * W + j;
            double ph = amp * orc_peaks(X, Y) + 2.0 * M_PI * carrier * (double)j;
            if (phi) phi[p] = ph;
            if (I) I[p] = 128.0 + 100.0 * cos(ph) + noise * orc_normal(seed, (uint64_t)t, p);
        }
    }
}

/* ========================================================================== */
/* WFT2 (WFF mode)                                                            */
/* ========================================================================== */
struct orc_wft2 {
    int W, H, PW, PH, sx, sy, nwx, nwy, nthreads, flags;
    double sigx, wxl, wxh, wxi, sigy, wyl, wyh, wyi, thr;
    double *gpad;       /* PH*PW complex: normalised window, top-left */
    double *gxn, *gyn;  /* 1-D normalised windows, length 2s+1 */
    double *fpad, *Ff;  /* PH*PW complex */
    /* per-thread */
    double *gwave, *Fg, *Sf, *filt;
    double *Gx, *Gy;    /* per-thread 1-D spectra (separable mode) */
    fft2ctx *ctx;
    fftplan *px, *py;
};

/* WFT2_cpu ctor + WFT2_Initialize + WFF2_Init, WFT/WFT2_CPU.cpp:76-121, 697-845 */
orc_wft2 *orc_wft2_create(int W, int H,
                          double sigx, double wxl, double wxh, double wxi,
                          double sigy, double wyl, double wyh, double wyi,
                          double thr, int nthreads, int flags) {
    if (W < 1 || H < 1 || nthreads < 1) return NULL;
    orc_wft2 *w = (orc_wft2 *)calloc(1, sizeof(orc_wft2));
    w->W = W; w->H = H;
    w->sigx = sigx; w->wxl = wxl; w->wxh = wxh; w->wxi = wxi;
    w->sigy = sigy; w->wyl = wyl; w->wyh = wyh; w->wyi = wyi;
    w->thr = thr;
    w->nthreads = nthreads;
    w->flags = flags;
    /* :701-706 */
    w->sx = (int)round(3 * sigx);
    w->sy = (int)round(3 * sigy);
    int winw = 2 * w->sx + 1, winh = 2 * w->sy + 1;
    /* :710-724 */
    int iH = orc_first_greater(H + winh - 1);
    int iW = orc_first_greater(W + winw - 1);
    if (iH < 0 || iW < 0) { free(w); return NULL; }   /* reference throws int -1 */
    w->PH = g_opt[iH];
    w->PW = g_opt[iW];
    w->nwx = orc_pair_count(wxl, wxh, wxi);
    w->nwy = orc_pair_count(wyl, wyh, wyi);
    size_t np = (size_t)w->PH * w->PW;
    w->gpad = (double *)calloc(2 * np, sizeof(double));
    w->fpad = (double *)calloc(2 * np, sizeof(double));
    w->Ff = (double *)calloc(2 * np, sizeof(double));
    /* window: :752-805 (exp of the summed exponent, scaled by 1/||g||_2) */
    double nrm = 0;
    for (int i = 0; i < winh; i++)
        for (int j = 0; j < winw; j++) {
            double y = (double)(i - (winh - 1) / 2);
            double x = (double)(j - (winw - 1) / 2);
            double g = exp(-(x * x) / 2 / sigx / sigx - (y * y) / 2 / sigy / sigy);
            w->gpad[2 * ((size_t)i * w->PW + j)] = g;
            nrm += g * g;
        }
    nrm = sqrt(nrm);
    for (int i = 0; i < winh; i++)
        for (int j = 0; j < winw; j++) w->gpad[2 * ((size_t)i * w->PW + j)] *= 1 / nrm;
    /* separable 1-D windows (same normalisation: ||g||_2 = ||gx||_2 ||gy||_2) */
    w->gxn = (double *)malloc(sizeof(double) * winw);
    w->gyn = (double *)malloc(sizeof(double) * winh);
    double nx = 0, ny = 0;
    for (int j = 0; j < winw; j++) {
        double x = (double)(j - w->sx);
        w->gxn[j] = exp(-(x * x) / 2 / sigx / sigx);
        nx += w->gxn[j] * w->gxn[j];
    }
    for (int i = 0; i < winh; i++) {
        double y = (double)(i - w->sy);
        w->gyn[i] = exp(-(y * y) / 2 / sigy / sigy);
        ny += w->gyn[i] * w->gyn[i];
    }
    nx = sqrt(nx); ny = sqrt(ny);
    for (int j = 0; j < winw; j++) w->gxn[j] /= nx;
    for (int i = 0; i < winh; i++) w->gyn[i] /= ny;
    /* per-thread buffers, :734-742, :836 */
    int T = nthreads;
    w->gwave = (double *)calloc(2 * np * T, sizeof(double));
    w->Fg = (double *)calloc(2 * np * T, sizeof(double));
    w->Sf = (double *)calloc(2 * np * T, sizeof(double));
    w->filt = (double *)calloc(2 * (size_t)W * H * T, sizeof(double));
    w->Gx = (double *)calloc(2 * (size_t)w->PW * T, sizeof(double));
    w->Gy = (double *)calloc(2 * (size_t)w->PH * T, sizeof(double));
    w->ctx = (fft2ctx *)malloc(sizeof(fft2ctx) * T);
    w->px = (fftplan *)malloc(sizeof(fftplan) * T);
    w->py = (fftplan *)malloc(sizeof(fftplan) * T);
    for (int t = 0; t < T; t++) {
        fft2_init(&w->ctx[t], w->PH, w->PW);
        plan_init(&w->px[t], w->PW);
        plan_init(&w->py[t], w->PH);
    }
    return w;
}

void orc_wft2_destroy(orc_wft2 *w) {
    if (!w) return;
    for (int t = 0; t < w->nthreads; t++) {
        fft2_free(&w->ctx[t]);
        plan_free(&w->px[t]);
        plan_free(&w->py[t]);
    }
    free(w->ctx); free(w->px); free(w->py);
    free(w->gpad); free(w->gxn); free(w->gyn); free(w->fpad); free(w->Ff);
    free(w->gwave); free(w->Fg); free(w->Sf); free(w->filt); free(w->Gx); free(w->Gy);
    free(w);
}

int orc_wft2_padded(const orc_wft2 *w, int *ph, int *pw) { *ph = w->PH; *pw = w->PW; return 0; }
int orc_wft2_halfwin(const orc_wft2 *w, int *sy, int *sx) { *sy = w->sy; *sx = w->sx; return 0; }
int orc_wft2_pairs(const orc_wft2 *w, int *nwy, int *nwx) { *nwy = w->nwy; *nwx = w->nwx; return 0; }
double orc_wft2_thr(const orc_wft2 *w) { return w->thr; }

/* WFF2_SetThreashold + feed, WFT/WFT2_CPU.cpp:959-1022 (thr cached after the
 * first call because m_rThr becomes non-negative). */
static void wff2_set_threshold_and_feed(orc_wft2 *w, const double *f) {
    double acc = 0;
    for (int i = 0; i < w->H; i++)
        for (int j = 0; j < w->W; j++) {
            size_t idf = (size_t)i * w->W + j, idp = (size_t)i * w->PW + j;
            double re = f[2 * idf], im = f[2 * idf + 1];
            if (w->thr < 0) {
                double a = sqrt(re * re + im * im);
                acc += a * a;
            }
            w->fpad[2 * idp] = re;
            w->fpad[2 * idp + 1] = im;
        }
    if (w->thr < 0) {
        acc = acc / (w->W * w->H) / 3;
        w->thr = 6 * sqrt(acc);
    }
}

/* 1-D spectrum of g1(k - s) * exp(j*wt*(k - s)) placed at [0, 2s] of length n */
static void spectrum_1d(const fftplan *p, const double *g1, int s, double wt,
                        double *out, double *work) {
    int n = p->n;
    memset(out, 0, sizeof(double) * 2 * (size_t)n);
    for (int k = 0; k <= 2 * s; k++) {
        double d = (double)(k - s);
        out[2 * k] = g1[k] * cos(wt * d);
        out[2 * k + 1] = g1[k] * sin(wt * d);
    }
    fft_exec(p, out, work, -1);
}

/* One (wy, wx) pair of WFF2 into the thread's accumulator acc,
 * WFT/WFT2_CPU.cpp:237-353: Fg, sf = ifft2(Ff .* Fg), crop + threshold,
 * fft2, .* Fg, ifft2, accumulate the crop.  Gy holds the pair's wy spectrum
 * (separable mode). */
static void wff2_pair(orc_wft2 *w, int tid, int y, int x, const double *Gy, double *acc) {
    const int W = w->W, H = w->H, PW = w->PW, PH = w->PH, sx = w->sx, sy = w->sy;
    const size_t np = (size_t)PH * PW;
    const double inv_np = 1.0 / (double)(PH * PW);
    const double thr = w->thr;
    double *gw = w->gwave + 2 * np * tid;
    double *Fg = w->Fg + 2 * np * tid;
    double *Sf = w->Sf + 2 * np * tid;
    double *Gx = w->Gx + 2 * (size_t)PW * tid;
    fft2ctx *ctx = &w->ctx[tid];
    double wyt = w->wyl + (double)y * w->wyi;
    double wxt = w->wxl + (double)x * w->wxi;
    if (w->flags & ORC_FG_SEPARABLE) {
        /* Fg(ky,kx) = Gy(ky) Gx(kx): separable form of FFT2(gwave) */
        spectrum_1d(&w->px[tid], w->gxn, sx, wxt, Gx, ctx->work);
        for (int i = 0; i < PH; i++)
            for (int j = 0; j < PW; j++) {
                size_t id = (size_t)i * PW + j;
                double ar = Gy[2 * i], ai = Gy[2 * i + 1];
                double br = Gx[2 * j], bi = Gx[2 * j + 1];
                Fg[2 * id] = ar * br - ai * bi;
                Fg[2 * id + 1] = ar * bi + ai * br;
            }
    } else {
        /* gwave = g .* exp(j(wxt*xx + wyt*yy)), :247-266; Fg = fft2(gwave), :268 */
        memset(gw, 0, sizeof(double) * 2 * np);
        for (int i = 0; i < 2 * sy + 1; i++)
            for (int j = 0; j < 2 * sx + 1; j++) {
                size_t id = (size_t)i * PW + j;
                double yy = i - sy, xx = j - sx;
                double c = cos(wxt * xx + wyt * yy), s = sin(wxt * xx + wyt * yy);
                double g = w->gpad[2 * id];
                gw[2 * id] = g * c;
                gw[2 * id + 1] = g * s;
            }
        memcpy(Fg, gw, sizeof(double) * 2 * np);
        fft2_exec(ctx, Fg, -1);
    }
    /* sf = ifft2(Ff .* Fg), :271-285 */
    for (size_t id = 0; id < np; id++) {
        double ar = w->Ff[2 * id], ai = w->Ff[2 * id + 1];
        double br = Fg[2 * id], bi = Fg[2 * id + 1];
        Sf[2 * id] = (ar * br - ai * bi) * inv_np;
        Sf[2 * id + 1] = (ar * bi + ai * br) * inv_np;
    }
    fft2_exec(ctx, Sf, +1);
    /* crop (+sy,+sx) and hard threshold |sf| >= thr, :289-319 */
    for (int i = 0; i < PH; i++)
        for (int j = 0; j < PW; j++) {
            size_t idp = (size_t)i * PW + j;
            if (i < H && j < W) {
                size_t idt = (size_t)(i + sy) * PW + j + sx;
                double re = Sf[2 * idt], im = Sf[2 * idt + 1];
                double a = sqrt(re * re + im * im);
                if (a >= thr) { Sf[2 * idp] = re; Sf[2 * idp + 1] = im; }
                else { Sf[2 * idp] = 0; Sf[2 * idp + 1] = 0; }
                Sf[2 * idt] = 0; Sf[2 * idt + 1] = 0;
            } else {
                Sf[2 * idp] = 0; Sf[2 * idp + 1] = 0;
            }
        }
    /* fft2, .* Fg / Np, ifft2, :322-339 */
    fft2_exec(ctx, Sf, -1);
    for (size_t id = 0; id < np; id++) {
        double ar = Sf[2 * id], ai = Sf[2 * id + 1];
        double br = Fg[2 * id], bi = Fg[2 * id + 1];
        Sf[2 * id] = (ar * br - ai * bi) * inv_np;
        Sf[2 * id + 1] = (ar * bi + ai * br) * inv_np;
    }
    fft2_exec(ctx, Sf, +1);
    /* accumulate the (+sy,+sx) crop, :341-352 */
    for (int i = 0; i < H; i++)
        for (int j = 0; j < W; j++) {
            size_t idf = (size_t)i * W + j;
            size_t ids = (size_t)(i + sy) * PW + (j + sx);
            acc[2 * idf] += Sf[2 * ids];
            acc[2 * idf + 1] += Sf[2 * ids + 1];
        }
}

/* WFF2, WFT/WFT2_CPU.cpp:190-390 (OpenMP over wy, thread-private buffers,
 * serial merge, final scale wxi*wyi/(4*pi^2)).  With ORC_PAIR_PARALLEL the
 * threads take (wy, wx) pairs instead of wy rows: the same arithmetic per
 * pair, more parallelism (the sum of the per-thread accumulators is in a
 * different order, a difference at FP64 rounding level). */
int orc_wft2_compute(orc_wft2 *w, const double *f, double *out) {
    const int W = w->W, H = w->H, PW = w->PW, PH = w->PH, sy = w->sy;
    const size_t np = (size_t)PH * PW, ni = (size_t)W * H;
    wff2_set_threshold_and_feed(w, f);
    memcpy(w->Ff, w->fpad, sizeof(double) * 2 * np);
    fft2_exec(&w->ctx[0], w->Ff, -1);
    memset(w->filt, 0, sizeof(double) * 2 * ni * w->nthreads);
    const int iwx = w->nwx, iwy = w->nwy;
#pragma omp parallel num_threads(w->nthreads)
    {
        int tid = 0, nth = 1;
#ifdef _OPENMP
        tid = omp_get_thread_num();
        nth = omp_get_num_threads();
#endif
        double *acc = w->filt + 2 * ni * tid;
        double *Gy = w->Gy + 2 * (size_t)PH * tid;
        fft2ctx *ctx = &w->ctx[tid];
        if (w->flags & ORC_PAIR_PARALLEL) {
            for (int p = tid; p < iwy * iwx; p += nth) {
                int y = p / iwx, x = p % iwx;
                if (w->flags & ORC_FG_SEPARABLE)
                    spectrum_1d(&w->py[tid], w->gyn, sy, w->wyl + (double)y * w->wyi, Gy,
                                ctx->work);
                wff2_pair(w, tid, y, x, Gy, acc);
            }
        } else {
            for (int y = tid; y < iwy; y += nth) {
                if (w->flags & ORC_FG_SEPARABLE)
                    spectrum_1d(&w->py[tid], w->gyn, sy, w->wyl + (double)y * w->wyi, Gy,
                                ctx->work);
                for (int x = 0; x < iwx; x++) wff2_pair(w, tid, y, x, Gy, acc);
            }
        }
    }
    /* serial merge :358-372 and scale :377-389 */
    memset(out, 0, sizeof(double) * 2 * ni);
    for (int k = 0; k < w->nthreads; k++) {
        const double *acc = w->filt + 2 * ni * k;
        for (size_t id = 0; id < ni; id++) {
            out[2 * id] += acc[2 * id];
            out[2 * id + 1] += acc[2 * id + 1];
        }
    }
    const double oneOverPi2 = 1.0 / (M_PI * M_PI);
    const double sc = 0.25 * oneOverPi2 * w->wxi * w->wyi;
    for (size_t id = 0; id < ni; id++) {
        out[2 * id] *= sc;
        out[2 * id + 1] *= sc;
    }
    return 0;
}

/* ========================================================================== */
/* DPRA                                                                       */
/* ========================================================================== */
struct orc_dpra {
    int W, H, rr, nthreads, flags;
    double *phiref, *phicurr, *dphiref, *dphi_frame;
    double *z, *filtered;
    orc_wft2 *wft;
};

orc_dpra *orc_dpra_create(const double *phi0, int W, int H, int rr,
                          double sigx, double wxl, double wxh, double wxi,
                          double sigy, double wyl, double wyh, double wyi,
                          double thr, int nthreads, int flags) {
    if (W < 1 || H < 1 || rr < 1 || nthreads < 1) return NULL;
    orc_wft2 *wft = orc_wft2_create(W, H, sigx, wxl, wxh, wxi, sigy, wyl, wyh, wyi, thr,
                                    nthreads,
                                    ORC_FG_SEPARABLE | (flags & ORC_DPRA_PAIR_PARALLEL
                                                        ? ORC_PAIR_PARALLEL : 0));
    if (!wft) return NULL;
    orc_dpra *d = (orc_dpra *)calloc(1, sizeof(orc_dpra));
    size_t n = (size_t)W * H;
    d->W = W; d->H = H; d->rr = rr; d->nthreads = nthreads; d->flags = flags;
    d->wft = wft;
    d->phiref = (double *)malloc(sizeof(double) * n);
    d->phicurr = (double *)calloc(n, sizeof(double));
    d->dphiref = (double *)calloc(n, sizeof(double));   /* dpra_cpu.cpp:19 */
    d->dphi_frame = (double *)calloc(n, sizeof(double));
    d->z = (double *)calloc(2 * n, sizeof(double));
    d->filtered = (double *)calloc(2 * n, sizeof(double));
    memcpy(d->phiref, phi0, sizeof(double) * n);         /* dpra_cpu.cpp:39-43 */
    return d;
}

void orc_dpra_destroy(orc_dpra *d) {
    if (!d) return;
    orc_wft2_destroy(d->wft);
    free(d->phiref); free(d->phicurr); free(d->dphiref); free(d->dphi_frame);
    free(d->z); free(d->filtered);
    free(d);
}

double orc_dpra_thr(const orc_dpra *d) { return d->wft->thr; }

/* Symmetric 3x3 solve standing in for LAPACKE_dsysv(COL_MAJOR,'U'),
 * DPRA/dpra_cpu.cpp:218-224.  Only the upper triangle of A is read.  A is PSD
 * by construction (A00 = 9 >= neighbour count), so det >= 0 in exact
 * arithmetic; det <= 1e-10 * A00*A11*A22 is treated as singular. */
int orc_solve3_sym(const double *A, double *b) {
    double m[3][4];
    m[0][0] = A[0]; m[0][1] = A[3]; m[0][2] = A[6];
    m[1][0] = A[3]; m[1][1] = A[4]; m[1][2] = A[7];
    m[2][0] = A[6]; m[2][1] = A[7]; m[2][2] = A[8];
    m[0][3] = b[0]; m[1][3] = b[1]; m[2][3] = b[2];
    double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1])
               - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0])
               + m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    double dp = fabs(m[0][0] * m[1][1] * m[2][2]);
    if (!(fabs(det) > 1e-10 * dp)) return 1;
    for (int k = 0; k < 3; k++) {
        int p = k;
        for (int i = k + 1; i < 3; i++)
            if (fabs(m[i][k]) > fabs(m[p][k])) p = i;
        if (p != k)
            for (int c = 0; c < 4; c++) { double t = m[k][c]; m[k][c] = m[p][c]; m[p][c] = t; }
        if (m[k][k] == 0) return 1;
        for (int i = k + 1; i < 3; i++) {
            double l = m[i][k] / m[k][k];
            for (int c = k; c < 4; c++) m[i][c] -= l * m[k][c];
        }
    }
    double x[3];
    for (int i = 2; i >= 0; i--) {
        double s = m[i][3];
        for (int c = i + 1; c < 3; c++) s -= m[i][c] * x[c];
        x[i] = s / m[i][i];
    }
    b[0] = x[0]; b[1] = x[1]; b[2] = x[2];
    return 0;
}

/* DPRA least squares, DPRA/dpra_cpu.cpp:153-232 (OpenMP row-cyclic).
 * Singular pixels: Delta = 0 (z = 1). */
int orc_dpra_ls(orc_dpra *d, const double *f, double *z) {
    const int W = d->W, H = d->H;
    const double *phr = d->phiref;
#pragma omp parallel num_threads(d->nthreads)
    {
        int tid = 0, nth = 1;
#ifdef _OPENMP
        tid = omp_get_thread_num();
        nth = omp_get_num_threads();
#endif
        double A[9], b[3];
        for (int i = tid; i < H; i += nth)
            for (int j = 0; j < W; j++) {
                double sc = 0, ss = 0, scs = 0, ss2 = 0, sc2 = 0, sf = 0, sfc = 0, sfs = 0;
                int m0 = i - 1 > 0 ? i - 1 : 0, m1 = i + 1 < H - 1 ? i + 1 : H - 1;
                int n0 = j - 1 > 0 ? j - 1 : 0, n1 = j + 1 < W - 1 ? j + 1 : W - 1;
                for (int m = m0; m <= m1; m++)
                    for (int n = n0; n <= n1; n++) {
                        size_t q = (size_t)m * W + n;
                        double c = cos(phr[q]), s = sin(phr[q]), ft = f[q];
                        sc += c; ss += s; scs += c * s; ss2 += s * s; sc2 += c * c;
                        sf += ft; sfc += ft * c; sfs += ft * s;
                    }
                A[0] = 9;   A[1] = 0;   A[2] = 0;
                A[3] = sc;  A[4] = sc2; A[5] = 0;
                A[6] = ss;  A[7] = scs; A[8] = ss2;
                b[0] = sf; b[1] = sfc; b[2] = sfs;
                size_t p = (size_t)i * W + j;
                if (orc_solve3_sym(A, b)) {
                    z[2 * p] = 1.0;
                    z[2 * p + 1] = 0.0;
                } else {
                    double dl = atan2(-b[2], b[1]);
                    z[2 * p] = cos(dl);
                    z[2 * p + 1] = sin(dl);
                }
            }
    }
    return 0;
}

/* dpra_per_frame, DPRA/dpra_cpu.cpp:130-252 */
int orc_dpra_frame(orc_dpra *d, const double *f, double *dphi, double *phicurr) {
    const size_t n = (size_t)d->W * d->H;
    orc_dpra_ls(d, f, d->z);
    orc_wft2_compute(d->wft, d->z, d->filtered);
    for (size_t i = 0; i < n; i++) {
        double t = atan2(d->filtered[2 * i + 1], d->filtered[2 * i]);
        if (!(d->flags & ORC_NO_FLOAT_CAST)) t = (double)(float)t;   /* :246 */
        d->phicurr[i] = atan2(sin(t + d->phiref[i]), cos(t + d->phiref[i]));
        d->dphi_frame[i] = d->dphiref[i] + t;
    }
    if (dphi) memcpy(dphi, d->dphi_frame, sizeof(double) * n);
    if (phicurr) memcpy(phicurr, d->phicurr, sizeof(double) * n);
    return 0;
}

/* update_ref_phi, DPRA/dpra_cpu.cpp:254-258 */
void orc_dpra_update_ref(orc_dpra *d) {
    const size_t n = (size_t)d->W * d->H;
    memcpy(d->phiref, d->phicurr, sizeof(double) * n);
    memcpy(d->dphiref, d->dphi_frame, sizeof(double) * n);
}

void orc_dpra_set_state(orc_dpra *d, const double *phiref, const double *dphiref) {
    const size_t n = (size_t)d->W * d->H;
    if (phiref) memcpy(d->phiref, phiref, sizeof(double) * n);
    if (dphiref) memcpy(d->dphiref, dphiref, sizeof(double) * n);
}

/* operator(), DPRA/dpra_cpu.cpp:58-89 */
int orc_dpra_run(orc_dpra *d, const double *frames, int T, double *dphi_all) {
    const size_t n = (size_t)d->W * d->H;
    for (int i = 0; i < T; i++) {
        orc_dpra_frame(d, frames + n * i, dphi_all ? dphi_all + n * i : NULL, NULL);
        if (i % d->rr == 0) orc_dpra_update_ref(d);
    }
    return 0;
}

/* ========================================================================== */
void orc_wrapped_err(const double *a, const double *b, int64_t n, double *rms, double *maxabs) {
    double s = 0, mx = 0;
    for (int64_t i = 0; i < n; i++) {
        double e = a[i] - b[i];
        e = atan2(sin(e), cos(e));
        s += e * e;
        if (fabs(e) > mx) mx = fabs(e);
    }
    *rms = n ? sqrt(s / (double)n) : 0;
    *maxabs = mx;
}

/* ========================================================================== */
/* AIA: advanced iterative algorithm for phase extraction from randomly        */
/* phase-shifted frames (Z. Wang, B. Han, Opt. Lett. 29, 1671 (2004)).        */
/* Restates MatLab/fp_aia.m:33-121, AIA/aia_cpu.cpp, AIA/aia_cpuf.cpp:16-260  */
/* and AIA/aia_cudaf.cu:290-461 in FP64.  Model I_k(j) = a_j + b_j cos(phi_j  */
/* + delta_k); each half-step is a 3x3 least squares in (a, b cos, -b sin).  */
/* ========================================================================== */

/* Cholesky A = U^T U of a symmetric 3x3 (upper triangle of the column-major
 * array read), as LAPACKE_?posv(.., 'U', ..) factors it (aia_cpuf.cpp:161,238).
 * Returns 0, or k (1..3) when the leading minor of order k is not positive
 * definite -- the reference prints that and exits (aia_cpuf.cpp:163-167).  A
 * pivot at or below 1e-10 of its diagonal entry counts as not positive (LAPACK
 * tests <= 0, but rounding can leave a singular matrix, e.g. all shifts
 * equal, a tiny positive pivot). */
#define ORC_CHOL_TOL 1e-10
static int chol3(const double *A, double U[3][3]) {
    const double a00 = A[0], a01 = A[3], a02 = A[6], a11 = A[4], a12 = A[7], a22 = A[8];
    memset(U, 0, 9 * sizeof(double));
    if (!(a00 > 0)) return 1;
    U[0][0] = sqrt(a00);
    U[0][1] = a01 / U[0][0];
    U[0][2] = a02 / U[0][0];
    double d1 = a11 - U[0][1] * U[0][1];
    if (!(d1 > ORC_CHOL_TOL * a11)) return 2;
    U[1][1] = sqrt(d1);
    U[1][2] = (a12 - U[0][1] * U[0][2]) / U[1][1];
    double d2 = a22 - U[0][2] * U[0][2] - U[1][2] * U[1][2];
    if (!(d2 > ORC_CHOL_TOL * a22)) return 3;
    U[2][2] = sqrt(d2);
    return 0;
}

static void chol3_solve(double U[3][3], double *b) {
    /* U^T y = b */
    double y0 = b[0] / U[0][0];
    double y1 = (b[1] - U[0][1] * y0) / U[1][1];
    double y2 = (b[2] - U[0][2] * y0 - U[1][2] * y1) / U[2][2];
    /* U x = y */
    double x2 = y2 / U[2][2];
    double x1 = (y1 - U[1][2] * x2) / U[1][1];
    double x0 = (y0 - U[0][1] * x1 - U[0][2] * x2) / U[0][0];
    b[0] = x0; b[1] = x1; b[2] = x2;
}

int orc_aia_chol3_solve(const double *A, double *b) {
    double U[3][3];
    int info = chol3(A, U);
    if (info) return info;
    chol3_solve(U, b);
    return 0;
}

/* computePhi, AIA/aia_cpuf.cpp:112-183 (MatLab/fp_aia.m:74-101):
 * A1 = sum_k v_k v_k^T with v_k = (1, cos delta_k, sin delta_k); per pixel
 * b = sum_k I_k v_k, x = A1^-1 b, phi = atan2(-x2, x1).  f: M x N frames
 * (frame-major).  X (optional): N x 3 solutions, as the reference's
 * v_b_phi after the solve (aia_cpuf.cpp:161, dumped to Test/Phi_CPU.csv by the
 * commented lines :169-178).  Returns 0 or the LAPACK-style info. */
int orc_aia_phi(const double *f, int M, int64_t N, const double *delta, double *phi,
                double *X, int nthreads) {
    double A[9] = {0};
    double *c = malloc(sizeof(double) * (size_t)M), *s = malloc(sizeof(double) * (size_t)M);
    for (int k = 0; k < M; k++) {
        c[k] = cos(delta[k]);
        s[k] = sin(delta[k]);
        A[3] += c[k]; A[4] += c[k] * c[k];
        A[6] += s[k]; A[7] += c[k] * s[k]; A[8] += s[k] * s[k];
    }
    A[0] = M;
    double U[3][3];
    int info = chol3(A, U);
    if (!info) {
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int64_t j = 0; j < N; j++) {
            double b[3] = {0, 0, 0};
            for (int k = 0; k < M; k++) {
                double I = f[(size_t)k * N + j];
                b[0] += I;
                b[1] += I * c[k];
                b[2] += I * s[k];
            }
            chol3_solve(U, b);
            if (X) { X[3 * j] = b[0]; X[3 * j + 1] = b[1]; X[3 * j + 2] = b[2]; }
            phi[j] = atan2(-b[2], b[1]);
        }
    }
    free(c);
    free(s);
    return info;
}

/* Pixel sums in a fixed order whatever the thread count: 256 static chunks,
 * each summed in sequence, then the chunks in order. */
#define ORC_AIA_CHUNKS 256

/* computeDelta, AIA/aia_cpuf.cpp:185-245 (MatLab/fp_aia.m:103-121):
 * A2 = sum_j u_j u_j^T with u_j = (1, cos phi_j, sin phi_j); per frame
 * b = sum_j I_k(j) u_j, x = A2^-1 b, delta_k = atan2(-x2, x1). */
int orc_aia_delta(const double *f, int M, int64_t N, const double *phi, double *delta,
                  int nthreads) {
    const int V = 5 + 3 * M;  /* sc, scc, ss, scs, sss, then (t, p, q) per frame */
    double *part = calloc((size_t)ORC_AIA_CHUNKS * V, sizeof(double));
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int ch = 0; ch < ORC_AIA_CHUNKS; ch++) {
        int64_t j0 = N * ch / ORC_AIA_CHUNKS, j1 = N * (ch + 1) / ORC_AIA_CHUNKS;
        double *acc = part + (size_t)ch * V;
        for (int64_t j = j0; j < j1; j++) {
            double cp = cos(phi[j]), sp = sin(phi[j]);
            acc[0] += cp; acc[1] += cp * cp; acc[2] += sp; acc[3] += cp * sp; acc[4] += sp * sp;
            for (int k = 0; k < M; k++) {
                double I = f[(size_t)k * N + j];
                acc[5 + 3 * k] += I;
                acc[6 + 3 * k] += I * cp;
                acc[7 + 3 * k] += I * sp;
            }
        }
    }
    double *tot = calloc((size_t)V, sizeof(double));
    for (int ch = 0; ch < ORC_AIA_CHUNKS; ch++)
        for (int v = 0; v < V; v++) tot[v] += part[(size_t)ch * V + v];
    double A[9] = {(double)N, 0, 0, tot[0], tot[1], 0, tot[2], tot[3], tot[4]};
    double U[3][3];
    int info = chol3(A, U);
    if (!info)
        for (int k = 0; k < M; k++) {
            double b[3] = {tot[5 + 3 * k], tot[6 + 3 * k], tot[7 + 3 * k]};
            chol3_solve(U, b);
            delta[k] = atan2(-b[2], b[1]);
        }
    free(part);
    free(tot);
    return info;
}

/* operator(), AIA/aia_cpuf.cpp:16-107 and AIA/aia_cudaf.cu:290-372: iterate
 * phi/delta estimation while err > max_err and iters < max_iter (err starts
 * at 2*max_err), one more phi estimate, then phi += delta_0 and
 * delta_k -= delta_0, both wrapped by atan2(sin, cos).  err_mode 0:
 * max_k |delta_k - delta_k_old| (aia_cudaf.cu:448-461, aia_cpu.cpp:249-263,
 * fp_aia.m:61); 1: max_k |(delta_k - delta_0) - (delta_k_old - delta_0_old)|
 * (aia_cpuf.cpp:247-260, aia_cudaf_v2.cu:665-676).  delta: in/out (M).
 * Returns 0 or the info of the failing factorisation. */
int orc_aia(const double *f, int M, int64_t N, double *delta, int max_iter, double max_err,
            int err_mode, double *phi, int *iters, double *err, int nthreads) {
    double *old = malloc(sizeof(double) * (size_t)M);
    int info = 0;
    *err = max_err * 2.0;
    *iters = 0;
    while (*err > max_err && *iters < max_iter) {
        memcpy(old, delta, sizeof(double) * (size_t)M);
        if ((info = orc_aia_phi(f, M, N, delta, phi, NULL, nthreads))) goto out;
        if ((info = orc_aia_delta(f, M, N, phi, delta, nthreads))) goto out;
        ++*iters;
        double e = 0;
        for (int k = 0; k < M; k++) {
            double d = err_mode ? fabs((delta[k] - delta[0]) - (old[k] - old[0]))
                                : fabs(delta[k] - old[k]);
            if (d > e) e = d;
        }
        *err = e;
    }
    if ((info = orc_aia_phi(f, M, N, delta, phi, NULL, nthreads))) goto out;
    {
        const double d0 = delta[0];
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int64_t j = 0; j < N; j++) {
            double t = phi[j] + d0;
            phi[j] = atan2(sin(t), cos(t));
        }
        for (int k = M - 1; k >= 0; k--) {
            double t = delta[k] - d0;
            delta[k] = atan2(sin(t), cos(t));
        }
    }
out:
    free(old);
    return info;
}

/* ========================================================================== */
/* WFR2: windowed Fourier ridges (the WFR mode of the same WFT2 classes).      */
/* ========================================================================== */

/* WFR2, WFT/WFT2_CPU.cpp:391-694 and MatLab/fp_wft2f.m:103-147, in FP64.
 * out: 8 planes of W*H doubles -- wx, wy, phase, phase_comp, b, r, cxx, cyy
 * (the members of WFT2_HostResults, WFT/WFT.h:52-60).
 *   ridge   for each pair (wy outer, wx inner) sf = f * gwave (the same
 *           convolution as WFF2), and where |sf| > r: r = |sf|, wx, wy and
 *           phase = angle(sf) are taken from that pair.  Strict > as in
 *           fp_wft2f.m:126 and WFT2_CUDAf.cu:637 (the C++ CPU class uses >=,
 *           WFT2_CPU.cpp:505, which differs only on exact ties: here the first
 *           pair in (wy, wx) order wins, and a pixel whose |sf| is 0 for every
 *           pair keeps r = wx = wy = phase = 0).
 *   curv    cxx = -conv2(wx_rep, x.*g)/sum(x.*x.*g), cyy likewise with y, where
 *           wx_rep replicates the border pixels by (sy, sx) (:540-628,
 *           fp_wft2f.m:133-141).  Evaluated directly and separably (g = gx gy,
 *           the normalisation cancels); the reference evaluates it by FFT over
 *           OPT(H+3sy) x OPT(W+3sx), which never wraps at the sampled points.
 *   comp    phase_comp = wrap(phase - atan(sx^2 cxx)/2 - atan(sy^2 cyy)/2),
 *           b = r ((1 + sigx^4 cxx^2)/(4 pi sigx^2))^(1/4) (same in y) (:674-691).
 * Returns -1 when OPT(H + 3sy) or OPT(W + 3sx) does not exist (WFR2_Init,
 * :848-858, where the reference constructor fails). */
int orc_wfr2_compute(orc_wft2 *w, const double *f, double *out) {
    const int W = w->W, H = w->H, PW = w->PW, PH = w->PH, sx = w->sx, sy = w->sy;
    const size_t np = (size_t)PH * PW, ni = (size_t)W * H;
    if (orc_first_greater(H + sy + 2 * sy) < 0 || orc_first_greater(W + sx + 2 * sx) < 0)
        return -1;
    /* feed (no threshold in WFR mode), Ff = fft2(fpad), :396-400 */
    for (int i = 0; i < H; i++)
        for (int j = 0; j < W; j++) {
            size_t idf = (size_t)i * W + j, idp = (size_t)i * PW + j;
            w->fpad[2 * idp] = f[2 * idf];
            w->fpad[2 * idp + 1] = f[2 * idf + 1];
        }
    memcpy(w->Ff, w->fpad, sizeof(double) * 2 * np);
    fft2_exec(&w->ctx[0], w->Ff, -1);
    const int iwx = w->nwx, iwy = w->nwy, T = w->nthreads;
    const double inv_np = 1.0 / (double)(PH * PW);
    /* per-thread ridge maps: r, pair order, sf (re, im) */
    double *rr = (double *)calloc(ni * T, sizeof(double));
    int *ord = (int *)malloc(sizeof(int) * ni * T);
    double *sre = (double *)calloc(ni * T, sizeof(double));
    double *sim = (double *)calloc(ni * T, sizeof(double));
    for (size_t i = 0; i < ni * T; i++) ord[i] = -1;
#pragma omp parallel num_threads(T)
    {
        int tid = 0, nth = 1;
#ifdef _OPENMP
        tid = omp_get_thread_num();
        nth = omp_get_num_threads();
#endif
        double *gw = w->gwave + 2 * np * tid;
        double *Fg = w->Fg + 2 * np * tid;
        double *Sf = w->Sf + 2 * np * tid;
        double *Gx = w->Gx + 2 * (size_t)PW * tid;
        double *Gy = w->Gy + 2 * (size_t)PH * tid;
        fft2ctx *ctx = &w->ctx[tid];
        double *r_t = rr + ni * tid, *re_t = sre + ni * tid, *im_t = sim + ni * tid;
        int *o_t = ord + ni * tid;
        for (int y = tid; y < iwy; y += nth) {
            double wyt = w->wyl + (double)y * w->wyi;
            if (w->flags & ORC_FG_SEPARABLE)
                spectrum_1d(&w->py[tid], w->gyn, sy, wyt, Gy, ctx->work);
            for (int x = 0; x < iwx; x++) {
                double wxt = w->wxl + (double)x * w->wxi;
                if (w->flags & ORC_FG_SEPARABLE) {
                    spectrum_1d(&w->px[tid], w->gxn, sx, wxt, Gx, ctx->work);
                    for (int i = 0; i < PH; i++)
                        for (int j = 0; j < PW; j++) {
                            size_t id = (size_t)i * PW + j;
                            double ar = Gy[2 * i], ai = Gy[2 * i + 1];
                            double br = Gx[2 * j], bi = Gx[2 * j + 1];
                            Fg[2 * id] = ar * br - ai * bi;
                            Fg[2 * id + 1] = ar * bi + ai * br;
                        }
                } else {
                    /* gwave = g .* exp(j(wxt*xx + wyt*yy)), :455-474 */
                    memset(gw, 0, sizeof(double) * 2 * np);
                    for (int i = 0; i < 2 * sy + 1; i++)
                        for (int j = 0; j < 2 * sx + 1; j++) {
                            size_t id = (size_t)i * PW + j;
                            double yy = i - sy, xx = j - sx;
                            double c = cos(wxt * xx + wyt * yy), s = sin(wxt * xx + wyt * yy);
                            double g = w->gpad[2 * id];
                            gw[2 * id] = g * c;
                            gw[2 * id + 1] = g * s;
                        }
                    memcpy(Fg, gw, sizeof(double) * 2 * np);
                    fft2_exec(ctx, Fg, -1);
                }
                /* sf = ifft2(Ff .* Fg), :476-493 */
                for (size_t id = 0; id < np; id++) {
                    double ar = w->Ff[2 * id], ai = w->Ff[2 * id + 1];
                    double br = Fg[2 * id], bi = Fg[2 * id + 1];
                    Sf[2 * id] = (ar * br - ai * bi) * inv_np;
                    Sf[2 * id + 1] = (ar * bi + ai * br) * inv_np;
                }
                fft2_exec(ctx, Sf, +1);
                /* ridge update on the (+sy, +sx) crop, :495-518 */
                const int order = y * iwx + x;
                for (int i = 0; i < H; i++)
                    for (int j = 0; j < W; j++) {
                        size_t idt = (size_t)(i + sy) * PW + j + sx, id = (size_t)i * W + j;
                        double re = Sf[2 * idt], im = Sf[2 * idt + 1];
                        double a = sqrt(re * re + im * im);
                        if (a > r_t[id]) {
                            r_t[id] = a;
                            o_t[id] = order;
                            re_t[id] = re;
                            im_t[id] = im;
                        }
                    }
            }
        }
    }
    /* merge (:523-540): largest r, ties to the earlier pair in (wy, wx) order */
    double *owx = out, *owy = out + ni, *oph = out + 2 * ni, *opc = out + 3 * ni;
    double *ob = out + 4 * ni, *orr = out + 5 * ni, *ocx = out + 6 * ni, *ocy = out + 7 * ni;
    for (size_t id = 0; id < ni; id++) {
        double best = 0;
        int bo = -1;
        double bre = 0, bim = 0;
        for (int t = 0; t < T; t++) {
            size_t k = ni * t + id;
            if (ord[k] >= 0 && (rr[k] > best || (rr[k] == best && bo >= 0 && ord[k] < bo))) {
                best = rr[k]; bo = ord[k]; bre = sre[k]; bim = sim[k];
            }
        }
        orr[id] = best;
        if (bo >= 0) {
            owx[id] = w->wxl + (double)(bo % iwx) * w->wxi;
            owy[id] = w->wyl + (double)(bo / iwx) * w->wyi;
            oph[id] = atan2(bim, bre);
        } else {
            owx[id] = owy[id] = oph[id] = 0;
        }
    }
    free(rr); free(ord); free(sre); free(sim);
    /* curvature: separable direct convolution with replicated borders */
    double *gx = (double *)malloc(sizeof(double) * (2 * sx + 1));
    double *gy = (double *)malloc(sizeof(double) * (2 * sy + 1));
    double sgx = 0, sgy = 0, sxxgx = 0, syygy = 0;
    for (int k = -sx; k <= sx; k++) {
        gx[k + sx] = exp(-(double)(k * k) / 2 / w->sigx / w->sigx);
        sgx += gx[k + sx];
        sxxgx += (double)k * k * gx[k + sx];
    }
    for (int k = -sy; k <= sy; k++) {
        gy[k + sy] = exp(-(double)(k * k) / 2 / w->sigy / w->sigy);
        sgy += gy[k + sy];
        syygy += (double)k * k * gy[k + sy];
    }
    double *rx = (double *)malloc(sizeof(double) * ni), *ry = (double *)malloc(sizeof(double) * ni);
#pragma omp parallel for num_threads(T) schedule(static)
    for (int i = 0; i < H; i++)
        for (int j = 0; j < W; j++) {
            double ax = 0, ay = 0;
            for (int k = -sx; k <= sx; k++) {
                int jj = j - k < 0 ? 0 : (j - k >= W ? W - 1 : j - k);
                ax += (double)k * gx[k + sx] * owx[(size_t)i * W + jj];
                ay += gx[k + sx] * owy[(size_t)i * W + jj];
            }
            rx[(size_t)i * W + j] = ax;
            ry[(size_t)i * W + j] = ay;
        }
    const double sgx2 = w->sigx * w->sigx, sgy2 = w->sigy * w->sigy;
#pragma omp parallel for num_threads(T) schedule(static)
    for (int i = 0; i < H; i++)
        for (int j = 0; j < W; j++) {
            double cx = 0, cy = 0;
            for (int k = -sy; k <= sy; k++) {
                int ii = i - k < 0 ? 0 : (i - k >= H ? H - 1 : i - k);
                cx += gy[k + sy] * rx[(size_t)ii * W + j];
                cy += (double)k * gy[k + sy] * ry[(size_t)ii * W + j];
            }
            size_t id = (size_t)i * W + j;
            cx = -cx / (sgy * sxxgx);
            cy = -cy / (sgx * syygy);
            ocx[id] = cx;
            ocy[id] = cy;
            double pc = oph[id] - 0.5 * atan(sgx2 * cx) - 0.5 * atan(sgy2 * cy);
            opc[id] = atan2(sin(pc), cos(pc));
            ob[id] = orr[id] * pow((1 + sgx2 * sgx2 * cx * cx) * 0.25 / M_PI / sgx2, 0.25) *
                     pow((1 + sgy2 * sgy2 * cy * cy) * 0.25 / M_PI / sgy2, 0.25);
        }
    free(rx); free(ry); free(gx); free(gy);
    return 0;
}
