/*
 * glsu_oracle.h -- FP64 CPU oracle for the G-LS3U per-frame hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference's
 * FP64 CPU path (DPRA/dpra_cpu.cpp + WFT/WFT2_CPU.cpp, WFF2 mode), used as the
 * correctness checker for the HIP library and as the timed CPU baseline.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product library (libglsu) never links or calls it.
 *
 * Every function cites the reference lines it restates.  Third-party arithmetic
 * in the reference (MKL's FFTW3 API and LAPACKE_dsysv) is replaced by the
 * oracle's own mixed-radix FP64 FFT and FP64 3x3 solver; see the header of
 * glsu_oracle.c for how each part is pinned.
 */
#ifndef GLSU_ORACLE_H
#define GLSU_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- OPT_FFT_SIZE table + getFirstGreater (WFT/WFT.h:18-32, WFT/WFT.cpp:10-23) */
int orc_opt_table(int *out, int cap);      /* returns table length                 */
int orc_first_greater(int x);              /* index into the table, or -1          */
int orc_opt_size(int x);                   /* table value at that index, or -1     */

/* ---- (wx,wy) grid size, WFT/WFT2_CPU.cpp:222-223 */
int orc_pair_count(double wl, double wh, double wi);

/* ---- FFT: unnormalised complex DFT, interleaved (re,im) doubles.
 *      sign = -1 forward (FFTW_FORWARD), +1 backward (FFTW_BACKWARD).      */
int  orc_fft1d(int n, double *data, int sign);
int  orc_fft2d(int rows, int cols, double *data, int sign);

/* ---- seeded synthetic fringe generator (shared definition with the GPU
 *      generator; counter-based PRNG + FP64 Box-Muller).                    */
uint64_t orc_hash3(uint64_t seed, uint64_t t, uint64_t p);
double   orc_normal(uint64_t seed, uint64_t t, uint64_t p);
double   orc_peaks(double X, double Y);
void orc_gen_phase(int W, int H, double amp, double carrier, double *phi);
void orc_gen_frame(int W, int H, int t, uint64_t seed, double amp0, double amp1,
                   double carrier, double noise, double *I, double *phi);

/* ---- WFT2 (WFF) ------------------------------------------------------------ */
typedef struct orc_wft2 orc_wft2;
enum { ORC_FG_SEPARABLE = 1 };   /* Fg as outer product of 1-D DFTs (exact form) */
/* Threads take (wy, wx) pairs instead of wy rows (the reference's loop,
 * WFT/WFT2_CPU.cpp:227-237, parallelises over wy only).  Same arithmetic per
 * pair; used where the oracle checks long sequences, not for the baseline. */
enum { ORC_PAIR_PARALLEL = 2 };

orc_wft2 *orc_wft2_create(int W, int H,
                          double sigx, double wxl, double wxh, double wxi,
                          double sigy, double wyl, double wyh, double wyi,
                          double thr, int nthreads, int flags);
void   orc_wft2_destroy(orc_wft2 *w);
int    orc_wft2_padded(const orc_wft2 *w, int *ph, int *pw);
int    orc_wft2_halfwin(const orc_wft2 *w, int *sy, int *sx);
int    orc_wft2_pairs(const orc_wft2 *w, int *nwy, int *nwx);
double orc_wft2_thr(const orc_wft2 *w);
/* f, out: H*W interleaved complex doubles.  Returns 0 or -1. */
int    orc_wft2_compute(orc_wft2 *w, const double *f, double *out);
/* WFR mode: out = 8 planes of W*H doubles (wx, wy, phase, phase_comp, b, r,
 * cxx, cyy).  Returns 0, or -1 when the curvature padding is out of range. */
int    orc_wfr2_compute(orc_wft2 *w, const double *f, double *out);

/* ---- DPRA (DPRA_CPU) ------------------------------------------------------ */
typedef struct orc_dpra orc_dpra;
enum { ORC_NO_FLOAT_CAST = 1 };  /* disable the (float) cast of dpra_cpu.cpp:246 */
enum { ORC_DPRA_PAIR_PARALLEL = 2 };  /* WFF2 with ORC_PAIR_PARALLEL */

orc_dpra *orc_dpra_create(const double *phi0, int W, int H, int rr,
                          double sigx, double wxl, double wxh, double wxi,
                          double sigy, double wyl, double wyh, double wyi,
                          double thr, int nthreads, int flags);
void orc_dpra_destroy(orc_dpra *d);
/* least-squares part only: z = exp(j*Delta) (interleaved complex). */
int  orc_dpra_ls(orc_dpra *d, const double *frame, double *z);
/* one frame: LS + WFF2 + phase update (dpra_per_frame).  phicurr may be NULL. */
int  orc_dpra_frame(orc_dpra *d, const double *frame, double *dphi, double *phicurr);
void orc_dpra_update_ref(orc_dpra *d);
/* Overwrite the reference phase and the cumulative phase change (H*W doubles
 * each; either may be NULL).  The reference has no such call: its state is
 * phi0 and zero at construction (DPRA/dpra_cpu.cpp:19, :39-43). */
void orc_dpra_set_state(orc_dpra *d, const double *phiref, const double *dphiref);
/* operator(): frames T*H*W, dphi_all T*H*W, with i % rr == 0 reference updates */
int  orc_dpra_run(orc_dpra *d, const double *frames, int T, double *dphi_all);
/* per-pixel solve used by DPRA (exposed for pin tests): A col-major 3x3 (only
 * upper triangle read), b[3] overwritten with x.  Returns 0, or 1 if singular. */
int  orc_solve3_sym(const double *A, double *b);
double orc_dpra_thr(const orc_dpra *d);

/* ---- AIA (AIA/aia_cpuf.cpp, AIA/aia_cudaf.cu, MatLab/fp_aia.m) ---------------
 * f: M frames of N pixels, frame-major.  Return 0, or k (1..3) when the leading
 * minor of order k of a normal matrix is not positive definite. */
int orc_aia_chol3_solve(const double *A, double *b);
int orc_aia_phi(const double *f, int M, int64_t N, const double *delta, double *phi,
                double *X, int nthreads);
int orc_aia_delta(const double *f, int M, int64_t N, const double *phi, double *delta,
                  int nthreads);
int orc_aia(const double *f, int M, int64_t N, double *delta, int max_iter, double max_err,
            int err_mode, double *phi, int *iters, double *err, int nthreads);

/* ---- metrics -------------------------------------------------------------- */
void orc_wrapped_err(const double *a, const double *b, int64_t n,
                     double *rms, double *maxabs);

#ifdef __cplusplus
}
#endif
#endif
