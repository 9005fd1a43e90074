/*
 * nudim_oracle.h -- FP64 CPU restatement of NUDIM's dual-space (HIO/ER) iteration.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * path (libnudim_hip.so).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never calls it.
 *
 * It restates, loop for loop, the numerics of the reference:
 *   src/grid.c:61-133           polar / uniform node tables
 *   src/reconstruction.c:304-733 projection FT, sample assembly, weights,
 *                               initial INFFT, HIO/ER loop, constraint
 *   src/denoise.c:67-249,320-378 TV, TV-height, 2D anisotropic diffusion
 *   src/projection_sim.c:282-407,516-608 2D1D simulator and its DFT mode
 * The third-party arithmetic the reference links (NFFT3 nfft_trafo/adjoint,
 * the NFFT3 CGNR solver, FFTW's unnormalised DFT) is restated from its
 * published definition: a Kaiser-Bessel NFFT with the reference's grid
 * n = 2*nextpow2(N) and 2m+2 = 14 taps, NFFT3's CGNR update order, and a
 * Stockham mixed-radix FFT.
 *
 * Complex arrays are interleaved (re, im) doubles, as fftw_complex.
 */
#ifndef NUDIM_ORACLE_H
#define NUDIM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- geometry (grid.c, reconstruction.c:372-426) ---- */
int  oracle_polar_nodes(double amin, double amax, double ainterval, int np, double *x);
int  oracle_uniform_nodes(int nx, int nz, double *x);
void oracle_pad_range(double gam_min, double interval, double *pad_min, double *pad_max);
void oracle_weights(int np, int nang, double padangle, double *w);
/* per node (index i < M): kind 0 unknown, 1 known, 2 held-out (0 deg); order = tilt file */
void oracle_mask(int np, int nang, double pad_min, double padangle, double gam_min,
                 double gam_max, double interval, int *kind, int *order);

/* ---- transforms ---- */
/* exact projection FT (reconstruction.c:306-370): proj[nmeas][nz][nx] -> tomo[nmeas][nz][np] c128 */
void oracle_projection_ft(int nx, int nz, int np, int nmeas, const double *proj, double *tomo);
/* direct NDFT, NFFT3 sign convention, square N: fhat[N][N] <-> f[M] */
void oracle_ndft_trafo(int N, int M, const double *x, const double *fhat, double *f);
void oracle_ndft_adjoint(int N, int M, const double *x, const double *f, double *fhat);
/* Kaiser-Bessel NFFT; N0 x N1 coefficients, n0 x n1 grid, W taps per dim */
typedef struct oracle_nfft oracle_nfft;
oracle_nfft *oracle_nfft_create(int N0, int N1, int n0, int n1, int W, int M, const double *x);
void oracle_nfft_destroy(oracle_nfft *p);
void oracle_nfft_trafo(oracle_nfft *p, const double *fhat, double *f);
void oracle_nfft_adjoint(oracle_nfft *p, const double *f, double *fhat);
double oracle_kb_beta(int W, double sigma);
/* unnormalised batched DFT along the slowest axis: data[len][howmany] c128, sign -1 fwd / +1 bwd */
void oracle_fft_strided(double *data, int len, int howmany, int sign);
void oracle_fft(double *data, int n, int sign);

/* ---- CGNR (NFFT3 solver, CGNR [| PRECOMPUTE_WEIGHT]) ---- */
/* y[M], f_hat_iter[N*N] in/out, w[M] or NULL; nsteps loop_one_step calls after before_loop */
void oracle_cgnr(oracle_nfft *p, const double *w, const double *y, double *fhat, int nsteps);

/* ---- regularisers and constraint (denoise.c, reconstruction.c:672-700) ---- */
/* volumes are [nz][ny][nx] (the reference img[z][y][x]) */
void oracle_tv_grad(const double *u, int nz, int ny, int nx, int ylo, int yhi, double *grad);
void oracle_tv(double *u, float lam, int niter, int nz, int ny, int nx);
void oracle_tv_height(double *u, float lam, int niter, int nz, int ny, int nx, int nheight, int ncenter);
int  oracle_diffusion2d(double *u, float lambda, float kappa, int niter, int nz, int ny, int nx);
void oracle_constraint(double *img, const double *img_last, int nz, int np, int nx, int is_hio,
                       int config_height, int nheight, int ncenter);

/* ---- end-to-end reconstruction (reconstruction.c:304-719) ---- */
typedef struct oracle_params {
  int nx, nz, np;
  double gam_min, gam_max, interval, padangle;
  int weight, denoise, method; /* method 0 DIFFUSION, 1 TV */
  int config_height, nheight, ncenter;
  int nhio, ner;
  int nthreads;
  int tv_full;            /* reconstruction-parallel.c:766: TV over the whole volume */
  float diffusion_lambda; /* 0: 0.25f; reconstruction-parallel.c:764 passes 0.4 */
} oracle_params;
/* proj[nmeas][nz][nx]; img0/img out [nz][np][np] (may be NULL); rtrace[nhio+ner] (may be NULL) */
int oracle_reconstruct(const oracle_params *p, const double *proj, double *img0, double *img,
                       double *rtrace);
/* one HIO/ER iteration on img (in/out) given assembled slices[nz][M] c128 and kind[M] */
int oracle_iterate(const oracle_params *p, double *img, double *img_last_out,
                   const double *slices, const int *kind, int nn, double *rfactor);
/* seconds for the per-plane Fourier step of planes [j0, j1) of one pass */
double oracle_time_fourier_planes(const oracle_params *p, const double *img, const double *slices,
                                  const int *kind, int j0, int j1);
/* initial INFFT from assembled slices -> img0 */
int oracle_initial(const oracle_params *p, const double *slices, double *img0);
/* assembled measured samples slices[nz][M] c128 and kind[M] from projections */
int oracle_assemble(const oracle_params *p, const double *proj, double *slices, int *kind);

/* ---- simulator (projection_sim.c) ---- */
/* density[nz][ny][nx] -> proj[nslices][nz][np] = creal(slice)/(np*nz), no noise */
int oracle_project_2d1d(int nx, int ny, int nz, int np, double gam_min, double gam_max,
                        double interval, const double *density, double *proj);
int oracle_project_dft(int nx, int ny, int nz, int np, double gam_min, double gam_max,
                       double interval, const double *density, double *proj);

void oracle_set_threads(int n);
double oracle_wtime(void);

#ifdef __cplusplus
}
#endif
#endif
