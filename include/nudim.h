/*
 * nudim.h -- C ABI of libnudim_hip.so, the MI355X (gfx950) implementation of
 * NUDIM's dual-space HIO/ER constraint iteration.
 *
 * The reference has no library API: its `reconstruction` main calls NFFT3,
 * FFTW3 and its own denoise.c directly.  Each entry point below names the
 * reference call (file:line) it replaces.  Conventions:
 *   - every function returns 0 on success and a negative NUDIM_E* code on
 *     failure; nudim_last_error() gives a thread-local message.  The library
 *     never aborts the host process.
 *   - the caller owns every host array; the library copies on load / get.
 *     The library owns all device memory through the opaque context.
 *   - complex arrays are interleaved (re, im) doubles, as fftw_complex.
 *   - volumes are double [nz][np][np] (the reference's img[z][y][x],
 *     src/reconstruction.c:527).
 *   - a context is used from one host thread at a time; distinct contexts
 *     may run on distinct threads and overlap on the device.
 */
#ifndef NUDIM_H
#define NUDIM_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { NUDIM_OK = 0, NUDIM_EINVAL = -1, NUDIM_ENOMEM = -2, NUDIM_EHIP = -3,
       NUDIM_ESTATE = -4, NUDIM_EBUDGET = -5, NUDIM_EAGAIN = -6 };

/* arithmetic of the device path */
enum { NUDIM_FP64_PLAIN = 0, /* FP64 everywhere, literal NFFT-based CGNR, reference NFFT sizes */
       NUDIM_FP32 = 1 };     /* FP32 data, FP64 tables/scalars, Toeplitz normal operator */

/* --method (src/reconstruction.c:44-48) */
enum { NUDIM_DIFFUSION = 0, NUDIM_TV = 1 };

/* CGNR normal-operator form */
enum { NUDIM_CG_AUTO = 0, NUDIM_CG_NFFT = 1, NUDIM_CG_TOEPLITZ = 2 };

typedef struct nudim_params {
  /* command-line surface of src/reconstruction.c:91-112 */
  int nx, nz, np;                     /* --nx, --nz, -p/--pad */
  double gam_min, gam_max, interval;  /* --min, --max, -t/--interval */
  double padangle;                    /* --padangle */
  int weight;                         /* --weight */
  int random;                         /* --random (unknown samples: uniform[0,10)) */
  unsigned long seed;                 /* MT19937 seed for --random (GSL default 0 -> 4357) */
  int config_height, nheight, ncenter;/* --height / --center */
  int denoise, method;                /* --denoise, --method */
  int nhio, ner;                      /* --nhio, --ner */
  /* device path */
  int precision;                      /* NUDIM_FP64_PLAIN or NUDIM_FP32 */
  int device;                         /* HIP device ordinal */
  int nfft_n;                         /* oversampled grid (0: reference 2*nextpow2(np) for FP64,
                                         autotuned for FP32) */
  int nfft_w;                         /* window taps per dimension (0: 14 for FP64, tuned FP32) */
  int cg_form;                        /* NUDIM_CG_* (0: NFFT for FP64, Toeplitz for FP32) */
  int use_graph;                      /* run passes as hipGraphs: a pass is captured once per
                                         starting state (volume buffers, HIO/ER) and replayed
                                         for later passes that start in it; passes that time
                                         kernels (nudim_profile) or dump run directly.  While a
                                         pass is being captured, synchronous copies of other
                                         threads of the process fail: use it where one context
                                         is driven at a time (the job server ignores it) */
  /* denoising variants of src/reconstruction-parallel.c (0 / 0 for the serial program) */
  int tv_full;                        /* TV over the whole volume even with a height band (:766) */
  float diffusion_lambda;             /* aniso_diffusion_2d lambda; 0: 0.25f (:764 passes 0.4) */
  unsigned int cu_mask[8];            /* CUs the job's kernels may use, bit i = CU i (as
                                         hipExtStreamCreateWithCUMask); all zero: every CU */
} nudim_params;

typedef struct nudim_ctx nudim_ctx;

const char *nudim_last_error(void);
void nudim_default_params(nudim_params *p);

/* ---- job lifecycle: replaces main() of src/reconstruction.c:50-733 ---- */

/* validation of src/reconstruction.c:245-302; geometry of :304-426 */
int nudim_create(const nudim_params *p, nudim_ctx **out);
void nudim_destroy(nudim_ctx *ctx);
/* number of tilt files nmeas and polar nodes M (src/reconstruction.c:304, 384-387) */
int nudim_sizes(const nudim_ctx *ctx, int *nmeas, int *M, int *nang);
/* projection FT + sample assembly, src/reconstruction.c:333-487;
 * proj is double [nmeas][nz][nx] (file j, row z, column x) */
int nudim_load_projections(nudim_ctx *ctx, const double *proj);
/* assembled samples directly: slices c128 [nz][M], kind[M] (0 unknown, 1 known, 2 held out) */
int nudim_load_samples(nudim_ctx *ctx, const double *slices, const int *kind);
/* initial INFFT + z IFFT, src/reconstruction.c:489-539 */
int nudim_initial(nudim_ctx *ctx);
/* `count` passes of the HIO/ER loop body, src/reconstruction.c:545-700.  The pass
 * counter nn continues across calls (HIO while nn < nhio).  rfactor[count] gets
 * R = sqrt(diff1/diff2) of each pass (:603-604), rfactor[count][nvol] for a
 * batch (R of every volume of each pass); may be NULL. */
int nudim_iterate(nudim_ctx *ctx, int count, double *rfactor);
/* current volume img[nz][np][np] (double) */
int nudim_get_volume(nudim_ctx *ctx, double *img);
/* replace the current volume (parity tests of single passes) and the pass counter */
int nudim_set_volume(nudim_ctx *ctx, const double *img, int nn);
/* at pass `nn` keep copies of img before and after denoising (the
 * grad_before/grad_after dump of src/reconstruction.c:637-658) */
int nudim_set_dump(nudim_ctx *ctx, int nn, double *before, double *after);

/* ---- batches: nvol volumes of one geometry reconstructed together ----
 * No reference counterpart: the reference runs one volume per process
 * (src/reconstruction.c:50-733).  A batch context runs the same passes as
 * nvol contexts with these parameters would, each volume with its own
 * projections, CGNR step sizes, denoising norms and R, but every kernel of a
 * pass covers all volumes (the Fourier side as nvol nz planes), so small
 * volumes fill the device.  Each volume's result equals its own job's to
 * rounding (bit for bit where the plane grouping of the reductions is the
 * same, e.g. nz a multiple of 64 or nvol odd).  Device memory is about nvol
 * times a single job's; a batch whose NFFT grid n^2 nz nvol reaches 2^31
 * elements is refused (NUDIM_EINVAL).  nudim_initial and nudim_iterate work on the whole
 * batch (rfactor[count][nvol]); nudim_load_projections, nudim_get_volume,
 * nudim_load_samples, nudim_set_volume and nudim_set_dump refuse a batch
 * (NUDIM_ESTATE).  A batch of one volume is a plain job. */
int nudim_batch_create(const nudim_params *p, int nvol, nudim_ctx **out);
/* projections of volume vol (0 <= vol < nvol), as nudim_load_projections;
 * nudim_initial needs every volume loaded */
int nudim_batch_load_projections(nudim_ctx *ctx, int vol, const double *proj);
/* current volume vol, as nudim_get_volume */
int nudim_batch_get_volume(nudim_ctx *ctx, int vol, double *img);
/* bytes of device memory the context holds */
size_t nudim_device_bytes(const nudim_ctx *ctx);
/* bytes a context with these parameters would hold (admission control) */
int nudim_estimate_bytes(const nudim_params *p, size_t *bytes);
/* NFFT grid size n and window width W a job with these parameters uses.  For
 * an FP32 job with nfft_n = nfft_w = 0 (and NUDIM_AUTOTUNE not "0") they are
 * measured on the device for its node set, once per process and node set: the
 * fastest pair among the fixed-length FFT sizes up to the default n whose
 * trafo and adjoint stay within 2e-6 relative L2 of the FP64 NFFT at the
 * reference's sizes.  That is, for a job whose passes run the NFFT (cg_form
 * NUDIM_CG_NFFT, or NUDIM_CG_START=nfft / NUDIM_R_FORM=nfft); every other
 * FP32 job runs its NFFT only at load and in the initial INFFT and always
 * uses the defaults n = smallest 2^a 3^b 5^c >= 2 np, W = 8, measured or
 * not.  nudim_create does the same; calling this first moves the
 * measurement out of the first job.  (No reference counterpart: the
 * reference fixes n = 2 nextpow2(np), 14 taps, src/reconstruction.c:393-405.) */
int nudim_autotune(const nudim_params *p, int *n, int *W);
/* kernel timing with HIP events on the job's stream during nudim_iterate:
 * class 0 off, 1 Toeplitz column pass (with the <p, T p> column sums),
 * 2 NFFT interpolation (both kernels of the patch form), 3 NFFT gather,
 * 4 other line FFT passes, 5 TV sweeps, 6 z-axis transforms, 7 Toeplitz row
 * forward pass (with the CG vector update p = z + beta p), 8 Toeplitz row
 * inverse pass (with z -= alpha q and the |z|^2 row sums), 9 two-input
 * column pass of the projected CG start, 10 CG vector update z -= alpha q
 * as a kernel of its own (only with NUDIM_CG_UPDATE=kernel), 11
 * anisotropic diffusion (one timed launch group per call: all 80 sweeps).
 * Selecting a
 * class resets the counters.  (No reference counterpart: the reference only
 * prints total wall time, src/reconstruction-parallel.c:881-883.) */
int nudim_profile(nudim_ctx *ctx, int kernel_class);
/* total device ms, launch count and algorithmic bytes of the timed launches */
int nudim_profile_read(nudim_ctx *ctx, double *ms, long *launches, double *bytes);
/* wait for all device work of the context */
int nudim_synchronize(nudim_ctx *ctx);
/* bandwidth probe (no reference counterpart; measurement only): the best of
 * `reps` passes of a plain 16-byte-vector kernel over `bytes` of device
 * memory on `device`, kind 0 read only, kind 1 copy; gbs = bytes moved
 * (read + written) per second / 1e9 */
int nudim_stream_probe(int device, size_t bytes, int kind, int reps, double *gbs);

/* ---- parity entry points (stateless; host arrays in the reference layouts) ---- */

/* nfft_trafo of the plan p1 (src/reconstruction.c:393-405, :571):
 * f[P][M] = sum_k fhat[P][N][N] exp(-2 pi i (k - N/2).x_m); x[M][2] as grid.c:71-72 */
int nudim_nfft_trafo(int precision, int N, int M, const double *x, int n, int W, int P,
                     const double *fhat, double *f);
/* nfft_adjoint (inside the NFFT3 solver): fhat[P][N][N] = sum_m f[P][M] exp(+...) */
int nudim_nfft_adjoint(int precision, int N, int M, const double *x, int n, int W, int P,
                       const double *f, double *fhat);
/* solver_before_loop_complex + nsteps x solver_loop_one_step_complex, CGNR
 * [| PRECOMPUTE_WEIGHT] (src/reconstruction.c:413, :506-510, :592-595);
 * y[P][M], fhat[P][N][N] in/out (initial guess), w[M] or NULL */
int nudim_cgnr(int precision, int form, int N, int M, const double *x, const double *w, int n,
               int W, int P, const double *y, double *fhat, int nsteps);
/* fftw_plan_many_dft(1,&len,howmany,data,NULL,howmany,1,...,sign) in place,
 * src/reconstruction.c:521-524, 556-559, 607-610; data c128 [len][howmany] */
int nudim_zfft(int precision, int len, int howmany, double *data, int sign);
/* total_variation_discrete(u, lam, niter, nz, ny, nx), src/denoise.c:67-151; u in/out */
int nudim_tv(int precision, double *u, int nz, int ny, int nx, float lam, int niter);
/* total_variation_discrete_height, src/denoise.c:153-249 */
int nudim_tv_height(int precision, double *u, int nz, int ny, int nx, float lam, int niter,
                    int nheight, int ncenter);
/* No counterpart in the reference: nudim_tv_height on caller DEVICE memory
 * that allocates nothing, for callers that cannot call hipMalloc (code under
 * HIP graph capture, tight serving loops, harnesses that forbid allocation).
 * d_u is a device array double [nz][ny][nx], in/out; the work is enqueued on
 * `stream` (a hipStream_t; NULL: the calling thread's default stream) on the
 * current device, and the call returns without waiting for it.  The sweeps'
 * buffers live in a file-scope device array of 68 MiB per GPU, held until the
 * process ends: 4 nz ny nx sizeof(T) + 2 x (ny 128 x 8) + 256 bytes, rounded
 * to 256-byte sections, must fit (FP64: about 2.2M voxels, 128^3 fits;
 * FP32: about 4.4M).  Larger volumes fail with NUDIM_EINVAL.  Calls on one GPU
 * share that array: calls that overlap on one GPU (other streams, other
 * threads) race. */
int nudim_tv_height_noalloc(int precision, double *d_u, int nz, int ny, int nx, float lam, int niter,
                            int nheight, int ncenter, void *stream);
/* aniso_diffusion_2d, src/denoise.c:320-378 (requires nz == nx, see :349-350) */
int nudim_diffusion2d(int precision, double *u, int nz, int ny, int nx, float lambda,
                      float kappa, int niter);
/* density constraint, src/reconstruction.c:672-700; img [nz][np][np] in/out */
int nudim_constraint(int precision, double *img, const double *img_last, int nz, int np, int nx,
                     int is_hio, int config_height, int nheight, int ncenter);
/* projection FT, src/reconstruction.c:306-367: proj [nmeas][nz][nx] -> tomo c128 [nmeas][nz][np] */
int nudim_projection_ft(int precision, int nx, int nz, int np, int nmeas, const double *proj,
                        double *tomo);
/* 2D1D simulator, src/projection_sim.c:282-405 without noise (FP64; nx == ny,
 * nx, nz, np even): density [nz][ny][nx] -> proj [nslices][nz][np] =
 * creal(slice)/(np*nz), nslices = (int)((gam_max-gam_min)/interval+0.5)+1 */
int nudim_project_2d1d(int nx, int ny, int nz, int np, double gam_min, double gam_max,
                       double interval, const double *density, double *proj);

/* ---- serving: continuous batching of reconstruction jobs ----
 * (No reference counterpart: the reference runs one job per process,
 * src/reconstruction.c main.)  Jobs are admitted by the server's policy
 * while their device memory (nudim_estimate_bytes) fits the budget; each
 * runs on its own stream and host slot thread, so jobs overlap on the
 * device.  Pending jobs with the same parameters as an admitted small one
 * (np^2 nz <= 2^21) join it as one batch (nudim_batch_create; up to
 * max_stack volumes, 2^23 np^2 nz in all, that fit the budget, in an odd
 * count unless nz is a multiple of 64), which takes one slot: larger volumes
 * fill the device alone.  A job's result does not depend on what else
 * runs. */
typedef struct nudim_server nudim_server;
enum { NUDIM_SERVE_FIFO = 0,      /* first pending job (arrival order) that fits */
       NUDIM_SERVE_SHORTEST = 1 }; /* least work (np^2 nz passes) that fits; a job waiting
                                      longer than max_wait_s goes first */
typedef struct {
  long long hbm_budget; /* bytes the admitted jobs may hold; <= 0: 90% of free device memory */
  int max_active;       /* jobs in flight at once (slot threads), >= 1 */
  int device;
  int policy;           /* NUDIM_SERVE_* */
  double max_wait_s;    /* aging bound of NUDIM_SERVE_SHORTEST; <= 0: 30 s */
  int small_cus;        /* > 0: CUs reserved for small jobs; a job of at most small_work runs on
                           the first small_cus CUs, every other job on the rest (jobs that set
                           their own cu_mask keep it), and small jobs have max_active slots of
                           their own */
  double small_work;    /* np^2 nz passes of the largest small job */
  int max_stack;        /* most jobs run as one batch; <= 0: 8; 1: every job alone */
} nudim_server_opts;
typedef struct {
  long long job_id;
  int status;                      /* NUDIM_OK or the failing call's code */
  double t_submit, t_start, t_end; /* seconds since the server was created */
  long long device_bytes;          /* the admission estimate */
  char error[256];
  int stacked;                     /* jobs of the batch it ran in (1: alone) */
} nudim_job_report;
int nudim_server_create(const nudim_server_opts *opts, nudim_server **out);
/* Copies proj [nmeas][nz][nx]; volume [nz][np][np] and rfactors [nhio+ner]
 * (or NULL) must stay valid until the job's report is returned.  A job whose
 * estimate exceeds the whole budget is refused with NUDIM_EBUDGET. */
int nudim_server_submit(nudim_server *s, const nudim_params *p, const double *proj, double *volume,
                        double *rfactors, long long *job_id);
/* the next finished job, in completion order; timeout_ms < 0 waits for one,
 * NUDIM_EAGAIN when none finished within timeout_ms */
int nudim_server_next(nudim_server *s, int timeout_ms, nudim_job_report *report);
int nudim_server_stats(nudim_server *s, int *pending, int *active, long long *bytes_in_use);
/* runs the pending and running jobs to completion, then frees the server
 * (reports not yet taken are dropped) */
void nudim_server_destroy(nudim_server *s);

#ifdef __cplusplus
}
#endif
#endif
