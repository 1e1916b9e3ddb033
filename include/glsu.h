/*
 * glsu.h -- C ABI of the MI355X (gfx950) G-LS3U hot path: DPRA least squares
 * + WFF2 windowed Fourier filtering + reference-phase update.
 *
 * Every entry point replaces a member of the reference's C++ classes (cited
 * per function, paths relative to the TWANG006/G-LS3U tree).  Plain pointers and
 * sizes only; device pointers are HIP device addresses, host pointers are
 * ordinary host memory.  All calls return GLSU_OK (0) or a negative status;
 * glsu_last_error() gives the message of the last failure on the calling
 * thread.  A handle is not re-entrant; different handles are independent.  All
 * work of a handle is ordered on its stream: the one passed at create time, or
 * the default (null) stream when NULL is passed, as in the reference's GPU
 * classes.
 */
#ifndef GLSU_H
#define GLSU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GLSU_OK 0
#define GLSU_ERR_ARG (-1)         /* invalid argument / null handle                 */
#define GLSU_ERR_PADDING (-2)     /* padded FFT size out of range (ref: throw -1)    */
#define GLSU_ERR_SIZE (-3)        /* frame size differs from the initialised size    */
#define GLSU_ERR_DEVICE (-4)      /* HIP runtime error                               */
#define GLSU_ERR_UNSUPPORTED (-5) /* configuration outside this library's limits      */
#define GLSU_ERR_NOT_PD (-6)      /* AIA normal matrix not positive definite (ref:
                                     prints and exits, AIA/aia_cpuf.cpp:163-167)    */

/* WFT_FPA::WFT::WFT_TYPE, Utils/WFT-FPA.h:72-76 */
#define GLSU_WFF 0
#define GLSU_WFR 1

/* frame element types accepted by the DPRA entry points */
#define GLSU_U8 0
#define GLSU_F32 1
#define GLSU_F64 2

/* WFF parameters of the WFT2 constructors, WFT/WFT2_CPU.h:30-36
 * (sigma, low, HIGH, step per axis, then the threshold; thr < 0 selects
 * 6*sqrt(mean|f|^2/3) computed on the first call and cached). */
typedef struct glsu_wff_params {
  double sigma_x, wxl, wxh, wxi;
  double sigma_y, wyl, wyh, wyi;
  double thr;
} glsu_wff_params;

/* Launch/plan choices picked by the autotuner (g-ls3u_amd/glsu/autotune.py).
 * Any field < 0 keeps the library default (ls_tile: also 0).  planx / plany:
 * FFT plan ids (glsu_fft_plans) along rows and columns; cols_waves:
 * wavefronts per workgroup of the column kernel (1, 2 or 4); rows_waves:
 * wavefronts per row of the row kernels (1..4); ls_tile: pixel tile of the
 * DPRA least-squares kernel, 1 = 64x8 (default), 2 = 64x4, 3 = 32x8,
 * 4 = 128x4, 5 = 64x16, 6 = 32x16 (batches always use 64x8).  Results do not
 * depend on the FFT plans beyond FP32 rounding and not at all on the launch
 * shapes.  No reference counterpart (the reference fixes its DPRA blocks,
 * DPRA/dpra_cudaf.cu:121-122). */
typedef struct glsu_tuning {
  int planx, plany, cols_waves, rows_waves, ls_tile;
} glsu_tuning;

typedef struct glsu_wft2 glsu_wft2;
typedef struct glsu_dpra glsu_dpra;

const char *glsu_last_error(void);
/* library version string */
const char *glsu_version(void);
/* number of visible HIP devices (0 when none) */
int glsu_device_count(int *count);

/* Defaults of the WFF-mode constructor WFT2_cpu(W, H, WFF, z, n),
 * WFT/WFT2_CPU.cpp:43-54 (sigma 10, [-2-3/sigma, 2+3/sigma] step 1/sigma, thr<0). */
int glsu_wff_params_default(glsu_wff_params *p);
/* Defaults of the WFR-mode constructor, WFT/WFT2_CPU.cpp:56-67 (sigma 10,
 * [-2, 2] step 0.025 on both axes; thr unused). */
int glsu_wfr_params_default(glsu_wff_params *p);
/* Parameters the reference DPRA constructors hard-code,
 * DPRA/dpra_cpu.cpp:23-25 and DPRA/dpra_cudaf.cu:116-118
 * (sigma 20, [-0.15f, 0.15f] step 0.05f, thr 15). */
int glsu_wff_params_dpra_reference(glsu_wff_params *p);

/* ---- WFT2 (WFF mode) --------------------------------------------------------
 * Replaces WFT2_CUDAF::WFT2_CUDAF(W, H, type, sx, wxl, wxh, wxi, sy, wyl, wyh,
 * wyi, thr, z, nStreams) + cuWFT2_Initialize + cuWFF2_Init,
 * WFT/WFT2_CUDAf.h:27-33, WFT/WFT2_CUDAf.cu:1321-1419 (and the FP64 CPU twin
 * WFT2_cpu, WFT/WFT2_CPU.h:30-36).  Returns GLSU_ERR_PADDING where the
 * reference constructor throws int -1 (padded size above 4096,
 * WFT/WFT2_CPU.cpp:69-73).  type GLSU_WFR selects the ridge mode
 * (glsu_wft2_compute_wfr). */
int glsu_wft2_create(glsu_wft2 **out, int width, int height, int type,
                     const glsu_wff_params *p, void *hip_stream);
/* Same, with explicit tuning (NULL = defaults); used by the autotuner.
 * No counterpart in the reference. */
int glsu_wft2_create_tuned(glsu_wft2 **out, int width, int height, int type,
                           const glsu_wff_params *p, void *hip_stream, const glsu_tuning *t);
/* Replaces WFT2_CUDAF::operator()(cufftComplex *d_f, WFT2_DeviceResultsF &d_z,
 * double &time), WFT/WFT2_CUDAf.h:38-40 -> cuWFF2, WFT/WFT2_CUDAf.cu:1016-1173.
 * d_f, d_filtered: device arrays of height*width complex64 (interleaved
 * re,im; row-major).  ms: if non-null, the call synchronises and returns the
 * elapsed device time in milliseconds; if null the call is asynchronous. */
int glsu_wft2_compute(glsu_wft2 *h, const void *d_f, void *d_filtered, double *ms);
/* Host-pointer form of the FP64/FP32 CPU functor operator(), WFT/WFT2_CPU.h:41-43:
 * copies f in, filters, copies the result out (PCIe-inclusive). */
int glsu_wft2_compute_host(glsu_wft2 *h, const float *h_f, float *h_filtered, double *ms);
/* info: padded FFT lengths (this library's, not the reference's OPT sizes),
 * half window sizes, grid sizes, and the threshold in effect (NaN until the
 * first call when thr < 0). */
/* The plan ids and launch shapes in use (after defaults and LDS limits are
 * applied).  No counterpart in the reference. */
int glsu_wft2_get_tuning(glsu_wft2 *h, glsu_tuning *t);

/* WFR mode (type GLSU_WFR at create): the ridge maps of WFT2_DeviceResultsF
 * (WFT/WFT.h:126-133), each height*width floats on the device. */
typedef struct glsu_wfr_results {
  float *wx, *wy, *phase, *phase_comp, *b, *r, *cxx, *cyy;
} glsu_wfr_results;
/* Replaces WFT2_CUDAF::operator() in WFR mode (cuWFR2, WFT/WFT2_CUDAf.cu:
 * 1174-1319) with the semantics of WFT2_cpu::WFR2 (WFT/WFT2_CPU.cpp:391-694)
 * and MatLab/fp_wft2f.m:103-147: per pixel the (wx, wy) pair of largest
 * |f * w_p| (first in (wy, wx) order on exact ties) gives r, wx, wy and phase;
 * cxx, cyy are the least-squares slopes of the border-replicated wx, wy maps;
 * phase_comp and b the compensated phase and amplitude.  The constructor
 * also fails (GLSU_ERR_PADDING) when W + 3 sx or H + 3 sy exceeds 4096
 * (WFR2_Init, :848-858). */
int glsu_wft2_compute_wfr(glsu_wft2 *h, const void *d_f, const glsu_wfr_results *z, double *ms);
/* Host form: h_out receives the eight maps in the order of glsu_wfr_results
 * (8 * height * width floats). */
int glsu_wft2_compute_wfr_host(glsu_wft2 *h, const float *h_f, float *h_out, double *ms);
/* Average device time of one kernel stage over reps launches, after one full
 * transform of d_f: 1 rows, 2 columns, 3 the rest (WFF: inverse rows; WFR:
 * ridge and curvature).  No reference counterpart (measurement). */
int glsu_wft2_time_kernel(glsu_wft2 *h, const void *d_f, int stage, int reps, double *ms);
int glsu_wft2_info(glsu_wft2 *h, int *nx, int *ny, int *sx, int *sy, int *nwx, int *nwy,
                   double *thr);
int glsu_wft2_destroy(glsu_wft2 *h);

/* One WFF-mode WFT2 (constructor + operator()) of a height*width complex64
 * field, d_f -> d_filtered (device, row-major), enqueued on hip_stream (NULL:
 * the null stream; the current device must be the stream's), with the plans
 * and launch shapes glsu_wft2_create picks by default.  It allocates nothing:
 * every intermediate (window spectra, twiddles, threshold, partial sums, the
 * row and column planes) lives in one file-scope device array of 96 MiB that
 * each device holds from the library's first use of it until the process
 * ends; the tables are computed on the device on every call.  thr < 0 is
 * computed from this call's field (nothing is cached between calls).  Fields
 * whose planes do not fit (2 * nwx * height * ceil8(width) * 8 bytes plus
 * under 1 MiB; e.g. 1024x1024 with up to 5 wx values) return
 * GLSU_ERR_UNSUPPORTED.  Calls that overlap on one device share the array and
 * race: order them on one stream (or with events).  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops). */
int glsu_wft2_compute_noalloc(int width, int height, const glsu_wff_params *p, const void *d_f,
                              void *d_filtered, void *hip_stream);

/* ---- DPRA ------------------------------------------------------------------
 * Replaces DPRA_CUDAF::DPRA_CUDAF(const float *Phi0, W, H, rr),
 * DPRA/dpra_cudaf.h:28-30, DPRA/dpra_cudaf.cu:96-164, and the FP64 CPU
 * DPRA_CPU(const double *Phi0, W, H, rr, nThreads), DPRA/dpra_cpu.h:30-33.
 * phi0: HOST array of height*width initial phase (row-major).  The WFF
 * parameters the reference hard-codes are exposed here; p == NULL selects the
 * reference DPRA's own (glsu_wff_params_dpra_reference).  rr: reference update
 * rate (>= 1). */
int glsu_dpra_create(glsu_dpra **out, const double *phi0, int width, int height, int rr,
                     const glsu_wff_params *p, void *hip_stream);
int glsu_dpra_create_f32(glsu_dpra **out, const float *phi0, int width, int height, int rr,
                         const glsu_wff_params *p, void *hip_stream);
/* Same, with explicit tuning (NULL = defaults).  No reference counterpart. */
int glsu_dpra_create_tuned(glsu_dpra **out, const double *phi0, int width, int height, int rr,
                           const glsu_wff_params *p, void *hip_stream, const glsu_tuning *t);
/* Replaces DPRA_CUDAF::dpra_per_frame(const cv::Mat &img, std::vector<float> &dPhi,
 * double &time), DPRA/dpra_cudaf.h:41-43, DPRA/dpra_cudaf.cu:207-284 (FP64 twin
 * DPRA/dpra_cpu.cpp:130-252).  d_frame: DEVICE array of height*width elements of
 * frame_dtype (GLSU_U8/F32/F64).  d_dphi: device float[height*width] receiving
 * the phase change w.r.t. frame 0 (may be NULL).  d_phicurr: device float
 * [height*width] receiving the wrapped current phase (may be NULL).  ms as in
 * glsu_wft2_compute. */
int glsu_dpra_frame(glsu_dpra *h, const void *d_frame, int frame_dtype, float *d_dphi,
                    float *d_phicurr, double *ms);
/* Host-pointer form, as the reference (memcpy to pinned, H2D, compute, D2H). */
int glsu_dpra_frame_host(glsu_dpra *h, const void *h_frame, int frame_dtype, float *h_dphi,
                         double *ms);
/* Replaces DPRA_CUDAF::update_ref_phi(), DPRA/dpra_cudaf.cu:286-290
 * (phiref <- phicurr, dphiref <- dphi). */
int glsu_dpra_update_ref(glsu_dpra *h);
/* Replaces DPRA_CPU::operator()(frames, dPhi_Sum, time), DPRA/dpra_cpu.cpp:58-89:
 * for i in [0, nframes): dpra_per_frame(frame i); if (i % rr == 0) update_ref.
 * d_frames: device, nframes*height*width elements; d_dphi_all: device float
 * [nframes*height*width] (may be NULL).  The reference's CUDA overloads of
 * operator() are empty stubs (DPRA/dpra_cudaf.cu:193-205). */
int glsu_dpra_run(glsu_dpra *h, const void *d_frames, int nframes, int frame_dtype,
                  float *d_dphi_all, double *ms);
/* Same chain replayed from captured HIP graphs (one per reference-buffer
 * parity).  Results equal glsu_dpra_run bit for bit.  No reference counterpart. */
int glsu_dpra_run_graph(glsu_dpra *h, const void *d_frames, int nframes, int frame_dtype,
                        float *d_dphi_all, double *ms);
/* Diagnostics: same chain as glsu_dpra_run with HIP events between the kernels
 * of every frame; ms[0..3] receive the summed device times (ms) of the
 * least-squares, row-forward, column and row-inverse+update kernels.
 * No reference counterpart (the reference prints per-step event times,
 * DPRA/dpra_cudaf.cu:219-283). */
int glsu_dpra_profile(glsu_dpra *h, const void *d_frames, int nframes, int frame_dtype,
                      float *d_dphi_all, double *ms);
/* Diagnostics: glsu_dpra_profile that also counts the column transforms the
 * column kernel skipped because a column's thresholded signal was zero on
 * every image row (exact: the skipped transform would add zero).  *skipped:
 * that count; *transforms: the count without skipping, nframes * nwx * width
 * * nwy.  No reference counterpart. */
int glsu_dpra_profile_skips(glsu_dpra *h, const void *d_frames, int nframes, int frame_dtype,
                            float *d_dphi_all, double *ms, unsigned long long *skipped,
                            unsigned long long *transforms);
/* Diagnostics: average device time (ms) of one launch of a WFF2 kernel, from
 * `reps` back-to-back launches between two HIP events on the handle's stream.
 * stage: 1 row-forward, 2 column, 3 row-inverse+update.  Re-runs the kernel on
 * the handle's current buffers (idempotent; call after at least one frame).
 * No reference counterpart. */
int glsu_dpra_time_kernel(glsu_dpra *h, int stage, int reps, double *ms);
/* State access for checkpoint/resume (the reference cannot seed dphiref,
 * DPRA/dpra_cpu.cpp:19).  Host float arrays of height*width; any may be NULL. */
int glsu_dpra_get_state(glsu_dpra *h, float *phiref, float *dphiref);
int glsu_dpra_set_state(glsu_dpra *h, const float *phiref, const float *dphiref);
/* copy the last pre-filter field z = exp(j*Delta) (complex64, height*width)
 * into the device array d_z, ordered on the handle's stream */
int glsu_dpra_copy_z(glsu_dpra *h, void *d_z);
/* the WFF threshold in effect (NaN before the first frame when thr < 0) */
int glsu_dpra_thr(glsu_dpra *h, double *thr);
/* The plan ids and launch shapes in use.  No counterpart in the reference. */
int glsu_dpra_get_tuning(glsu_dpra *h, glsu_tuning *t);
int glsu_dpra_destroy(glsu_dpra *h);

/* ---- many sequences at once ----------------------------------------------
 * A batch runs the DPRA chain of n independent sequences of any sizes
 * together: per frame, one persistent least-squares launch over all images
 * and, per group of sequences sharing FFT plans and launch shapes, one
 * persistent launch per WFF2 stage, replayed from one captured hipGraph.
 * Each sequence behaves as its own DPRA_CPU / DPRA_CUDAF object
 * (DPRA/dpra_cpu.h:30-48: ctor(Phi0, W, H, rr), operator()(frames, dPhi_Sum,
 * time), update every frame with i % rr == 0) and gives bit-identical results
 * to glsu_dpra_run on a handle created with the same tuning.  The reference
 * has no batched form; its application runs one sequence at a time
 * (App_DPRA/dprawidget.cpp:223-264). */
typedef struct glsu_seq_desc {
  int width, height, rr;
  const double *phi0;      /* host, height*width initial phase */
} glsu_seq_desc;
typedef struct glsu_dpra_batch glsu_dpra_batch;
/* tunings: n entries (one per sequence) or NULL for the defaults. */
int glsu_dpra_batch_create(glsu_dpra_batch **out, const glsu_seq_desc *seqs, int n,
                           const glsu_wff_params *p, void *hip_stream,
                           const glsu_tuning *tunings);
/* d_frames[i]: device frames of sequence i (nframes*height*width, frame_dtype);
 * d_dphi_all: NULL or n device outputs (nframes*height*width float32).
 * Every call processes frames 0..nframes-1 of every sequence, continuing
 * from the state the previous call left (as repeated operator() calls). */
int glsu_dpra_batch_run(glsu_dpra_batch *h, const void *const *d_frames, int nframes,
                        int frame_dtype, float *const *d_dphi_all, int use_graph, double *ms);
int glsu_dpra_batch_info(glsu_dpra_batch *h, int *nclasses, int *ntiles);
int glsu_dpra_batch_get_tuning(glsu_dpra_batch *h, int seq, glsu_tuning *t);
int glsu_dpra_batch_get_state(glsu_dpra_batch *h, int seq, float *phiref, float *dphiref);
int glsu_dpra_batch_destroy(glsu_dpra_batch *h);

/* ---- learned WFF2 surrogate ---------------------------------------------
 * A U-Net (levels x base channels; 3x3 convs, relu, max pooling, nearest x2
 * upsampling with skip concatenation, 1x1 output with a residual on the
 * input) that maps the raw DPRA field z = (cos, sin) of the phase difference
 * to the WFF2-filtered unit vector.  BF16 weights and activations on MFMA
 * with FP32 accumulation; FP32 master weights and Adam.  No counterpart in
 * the reference: it stands in for WFT2_cpu::operator() (WFT/WFT2_CPU.cpp:
 * 178-186) inside dpra_per_frame, and WFF2 stays its fallback and arbiter.
 * Parameters are flat FP32 in layer order (e0a e0b ... e{L-1}b d{L-2}a d{L-2}b
 * ... d0a d0b out), each layer's weights [Cout][k][k][Cin] then biases.
 * Images: n x height x width complex64, height a multiple of 4*2^(levels-1),
 * width of 2^(levels-1). */
typedef struct glsu_unet glsu_unet;
int glsu_unet_create(glsu_unet **out, int levels, int base, unsigned long long seed,
                     void *hip_stream);
int glsu_unet_param_count(glsu_unet *h, long *count);
int glsu_unet_get_params(glsu_unet *h, float *host);
int glsu_unet_set_params(glsu_unet *h, const float *host);
int glsu_unet_get_grads(glsu_unet *h, float *host);
int glsu_unet_forward(glsu_unet *h, const void *d_x, void *d_out, int n, int height, int width);
/* Loss L = mean |out - u|^2 (u: target unit vectors, complex64) and its
 * gradients into the handle (read them with glsu_unet_get_grads). */
int glsu_unet_loss_grad(glsu_unet *h, const void *d_x, const void *d_u, int n, int height,
                        int width, double *loss);
int glsu_unet_adam_step(glsu_unet *h, float lr, float beta1, float beta2, float eps);
int glsu_unet_destroy(glsu_unet *h);
/* dpra_per_frame with the network in place of WFF2 (DPRA/dpra_cpu.cpp:238),
 * and a WFF2 spot check on the 64x64 tile at (tile_x, tile_y): when the RMS
 * wrapped difference between the network's phase and WFF2's exceeds tol
 * (radians), the frame is filtered by WFF2 instead (*fallback = 1).  tol < 0
 * disables the check.  *residual receives the RMS (or -1). */
int glsu_dpra_frame_learned(glsu_dpra *h, glsu_unet *net, const void *d_frame, int frame_dtype,
                            float *d_dphi, double tol, int tile_x, int tile_y, double *residual,
                            int *fallback);
/* With residual and fallback both NULL the call is asynchronous: the spot
 * check's decision stays on the device and selects which of the two enqueued
 * paths runs.  Spot-check statistics since the last reset (no counterpart in
 * the reference): checked frames, fallbacks, mean and max tile RMS. */
int glsu_dpra_learned_stats(glsu_dpra *h, int *frames, int *fallbacks, double *mean_rms,
                            double *max_rms, int reset);
/* The network output (complex64, height*width) of the last learned frame. */
int glsu_dpra_copy_net(glsu_dpra *h, void *d_out);

/* ---- AIA: initial phase from randomly phase-shifted frames -----------------
 * Replaces AIA::AIA_CUDAF (AIA/aia_cudaf.h:15-72): the constructor
 * AIA_CUDAF(iM, icols, irows) (:23-25, AIA/aia_cudaf.cu:248-285) and
 * operator() (:28-38, AIA/aia_cudaf.cu:290-372); the same algorithm as
 * AIA_CPU_Dn / AIA_CPU_DnF (AIA/aia_cpu.h:10, AIA/aia_cpuf.h:15) and
 * MatLab/fp_aia.m.  nframes: 3..1024. */
typedef struct glsu_aia glsu_aia;
int glsu_aia_create(glsu_aia **out, int nframes, int width, int height, void *hip_stream);
/* d_frames: nframes*height*width elements of frame_dtype, frame-major (device).
 * deltas (host, nframes): in, the initial shifts when use_given != 0, else
 * delta_0 = 0 and N(0,1) draws from `seed` (the reference draws them from a
 * clock seed when v_deltas.size() != M, AIA/aia_cpuf.cpp:31-45); out, the
 * final shifts minus delta_0, wrapped (AIA/aia_cudaf.cu:364-371).
 * Iterates while err > max_err and iters < max_iter (err starts at
 * 2*max_err).  err_mode 0: err = max_k |delta_k - delta_k_old| (AIA_CUDAF,
 * aia_cudaf.cu:448-461; AIA_CPU_Dn; fp_aia.m:61); 1: the same on shifts
 * relative to delta_0 (AIA_CPU_DnF, aia_cpuf.cpp:247-260; AIA_CUDAF_YC2).
 * d_phi (device, height*width floats): phi + delta_0 wrapped into (-pi, pi]
 * (aia_cudaf.cu:351-356).  *ms: GPU time of the whole solve.  Returns
 * GLSU_ERR_NOT_PD when a normal matrix is not positive definite. */
int glsu_aia_run(glsu_aia *h, const void *d_frames, int frame_dtype, float *deltas, int use_given,
                 uint64_t seed, int max_iter, double max_err, int err_mode, float *d_phi,
                 int *iters, double *err, double *ms);
/* The same from host frames (one pointer per frame, as the reference copies
 * each cv::Mat, aia_cudaf.cu:320-324) to a host phase map. */
int glsu_aia_run_host(glsu_aia *h, const void *const *h_frames, int frame_dtype, float *deltas,
                      int use_given, uint64_t seed, int max_iter, double max_err, int err_mode,
                      float *h_phi, int *iters, double *err, double *ms);
/* Average GPU time over reps repetitions, with the convergence test off, from
 * the given initial shifts, of one iteration (stage 0: the pass launch that
 * does the phase step and the shift-step sums, plus the one-block solve) or of
 * the pass launch alone (stage 1).  No reference counterpart (measurement). */
int glsu_aia_time_iter(glsu_aia *h, const void *d_frames, int frame_dtype, const float *deltas,
                       int stage, int reps, double *ms);
/* Launch shape: blocks per frame chunk, chunks, frames per chunk. */
int glsu_aia_info(glsu_aia *h, int *blocks, int *chunks, int *frames_per_chunk);
int glsu_aia_destroy(glsu_aia *h);

/* ---- the reference's matrix files ------------------------------------------
 * Binary matrix, replaces ReadMatrixFromDisk / WriteMatrixToDisk<T>
 * (Utils/matrixIO.inl:33-62, 77-104; header Utils/matrixIO.h:18-35): int32
 * rows, int32 cols, rows*cols elements of elem_size bytes, row-major.  With
 * data == NULL the read only returns the size; capacity is in bytes. */
int glsu_read_matrix_bin(const char *path, int elem_size, int *rows, int *cols, void *data,
                         size_t capacity);
int glsu_write_matrix_bin(const char *path, int elem_size, int rows, int cols, const void *data);
/* Complex text, replaces fftwComplexMatRead2D / fftwComplexMatWrite2D
 * (Utils/Utils.cpp:80-193): "rows,cols" then "re,im" per element.  data:
 * interleaved doubles; capacity in doubles; precision <= 0 keeps the stream
 * default of the reference (6 significant digits). */
int glsu_read_complex_text(const char *path, int *rows, int *cols, double *data,
                           size_t capacity);
int glsu_write_complex_text(const char *path, int rows, int cols, const double *data,
                            int precision);

/* ---- output stage ----------------------------------------------------------
 * Phase (e.g. the cumulative dphi of the chain, device floats) to 8-bit video
 * frames as the reference demo writes them (App_DPRA/dprawidget.cpp:541-549):
 * uchar((atan2(sin p, cos p) + pi) / 2 / pi * 255) in FP32, truncated.  n
 * pixels (any number of frames), asynchronous on hip_stream (NULL: default). */
int glsu_phase_to_u8(const float *d_phi, uint8_t *d_out, size_t n, void *hip_stream);

/* ---- utilities ---------------------------------------------------------- */
/* Seeded synthetic fringe frames, written to device memory (float32,
 * nframes*height*width): I = 128 + 100 cos(phi) + noise*N(0,1),
 * phi = (amp0 + amp1*t) peaks(X,Y) + 2*pi*carrier*x.  No reference counterpart
 * (the reference reads image files, App_DPRA/dprawidget.cpp:243-255). */
int glsu_gen_frames(float *d_out, int width, int height, int t0, int nframes, uint64_t seed,
                    double amp0, double amp1, double carrier, double noise, void *hip_stream);
/* FFT plans compiled into the library: fills up to cap entries of
 * (id, T, E, N) and returns the count in *count. */
int glsu_fft_plans(int *ids, int *lanes, int *elems, int *lengths, int cap, int *count);
/* How FFT plan `id` moves data between the lanes of a transform: 0 = every
 * cross-lane stage through LDS; 1 = the transition into the first stage (and
 * out of it in the inverse) through registers (v_permlane32/16_swap between
 * the wavefront's 16-lane rows), the other stages through LDS.  The kind
 * changes no arithmetic.  Returns the kind, or GLSU_ERR_ARG for an unknown
 * id.  No reference counterpart (plan choice is internal to this library). */
int glsu_fft_plan_exchange(int id);

#ifdef __cplusplus
}
#endif
#endif /* GLSU_H */
