/*
 * nudim_oracle.c -- FP64 CPU restatement of NUDIM's HIO/ER iteration.
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker and the CPU baseline).  The
 * product library never links or calls this file.  Each function names the
 * reference lines it follows.  See nudim_oracle.h for the list.
 */
#include "nudim_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* grid.h:11 -- the reference's truncated pi, used for node angles and weights */
#define REF_PI 3.1415926
#define TWO_PI 6.283185307179586476925286766559

typedef double complex cplx;

double oracle_wtime(void) {
  struct timeval tv;
  gettimeofday(&tv, NULL);
  return tv.tv_sec + 1e-6 * tv.tv_usec;
}

void oracle_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ------------------------------------------------------------------------ */
/* geometry                                                                  */
/* ------------------------------------------------------------------------ */

/* grid.c:61-97 (grid_slice_2D, equal_angle_sample_2D): accumulated angle and
 * radius; x[2i] = r sin(gamma/180*PI), x[2i+1] = r cos(gamma/180*PI). */
int oracle_polar_nodes(double amin, double amax, double ainterval, int np, double *x) {
  int nslices = (int)((amax - amin) / ainterval + 0.5) + 1;
  if (x) {
    double angle = amin - ainterval;
    for (int s = 0; s < nslices; s++) {
      angle += ainterval;
      double interval = 1.0 / np;
      double x0 = -0.5 - interval;
      double *xs = x + (size_t)s * np * 2;
      for (int i = 0; i < np; i++) {
        x0 += interval;
        xs[2 * i + 0] = x0 * sin(angle / 180 * REF_PI);
        xs[2 * i + 1] = x0 * cos(angle / 180 * REF_PI);
      }
    }
  }
  return nslices * np;
}

/* grid.c:100-133 (uniform_sample_2D): X fastest, accumulated. */
int oracle_uniform_nodes(int nx, int nz, double *x) {
  double intervalx = 1.0 / nx, intervalz = 1.0 / nz;
  double z0 = -0.5 - intervalz;
  for (int j = 0; j < nz; j++) {
    z0 += intervalz;
    double x0 = -0.5 - intervalx;
    for (int i = 0; i < nx; i++) {
      x0 += intervalx;
      x[2 * (j * nx + i) + 0] = x0;
      x[2 * (j * nx + i) + 1] = z0;
    }
  }
  return nx * nz;
}

/* reconstruction.c:372-379 */
void oracle_pad_range(double gam_min, double interval, double *pad_min, double *pad_max) {
  int ntmp = (int)((-90 - gam_min) / interval);
  *pad_min = gam_min + ntmp * interval;
  ntmp = (int)((90 - gam_min) / interval);
  *pad_max = gam_min + ntmp * interval;
}

/* reconstruction.c:415-426: integer abs and integer np/2 */
void oracle_weights(int np, int nang, double padangle, double *w) {
  for (int j = 0; j < nang; j++)
    for (int i = 0; i < np; i++) {
      if (i == np / 2)
        w[j * np + i] = REF_PI / (4 * nang * pow(np / 2, 2));
      else
        w[j * np + i] = padangle / 180 * REF_PI * abs(i - np / 2) / pow(np / 2, 2);
    }
}

/* reconstruction.c:437-484 (the test does not depend on the plane j) */
void oracle_mask(int np, int nang, double pad_min, double padangle, double gam_min,
                 double gam_max, double interval, int *kind, int *order) {
  int M = nang * np;
  for (int i = 0; i < M; i++) {
    int nslc = i / np;
    double angle = pad_min + nslc * padangle;
    int span = (int)(interval / padangle + 0.5);
    int seq = (int)((angle - gam_min) / padangle + 0.5);
    order[i] = -1;
    if ((angle < gam_min) || (angle > gam_max) || (seq % span != 0)) {
      kind[i] = 0;
    } else if (abs((int)angle) < 1e-5) {
      kind[i] = 2;
    } else {
      kind[i] = 1;
      order[i] = (int)((angle - gam_min) / interval + 0.5);
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Stockham mixed-radix FFT (unnormalised; sign -1 forward, +1 backward)     */
/* ------------------------------------------------------------------------ */

#define OFFT_MAXR 64
typedef struct {
  int n, nf, f[48];
  cplx *tw; /* tw[j] = exp(-2 pi i j / n) */
} offt_plan;

static offt_plan *offt_create(int n) {
  offt_plan *P = (offt_plan *)calloc(1, sizeof(offt_plan));
  P->n = n;
  int m = n;
  while (m % 4 == 0) { P->f[P->nf++] = 4; m /= 4; }
  while (m % 2 == 0) { P->f[P->nf++] = 2; m /= 2; }
  for (int r = 3; m > 1; r += 2)
    while (m % r == 0) { P->f[P->nf++] = r; m /= r; }
  P->tw = (cplx *)malloc(sizeof(cplx) * (n > 0 ? n : 1));
  for (int j = 0; j < n; j++) {
    double a = -TWO_PI * (double)j / (double)n;
    P->tw[j] = cos(a) + I * sin(a);
  }
  return P;
}

static void offt_destroy(offt_plan *P) {
  if (!P) return;
  free(P->tw);
  free(P);
}

/* one transform of x[n] (contiguous); work[n] scratch */
static void offt_exec(const offt_plan *P, cplx *x, cplx *work, int sign) {
  int n = P->n;
  cplx *src = x, *dst = work;
  int l = 1;
  for (int s = 0; s < P->nf; s++) {
    int r = P->f[s], lr = l * r, m = n / lr;
    for (int k = 0; k < l; k++) {
      for (int q = 0; q < m; q++) {
        cplx a[OFFT_MAXR];
        for (int t = 0; t < r; t++) {
          cplx w = P->tw[t * k * m];
          if (sign > 0) w = conj(w);
          a[t] = src[(k * r + t) * m + q] * w;
        }
        if (r == 2) {
          dst[k * m + q] = a[0] + a[1];
          dst[(k + l) * m + q] = a[0] - a[1];
        } else if (r == 4) {
          cplx s02 = a[0] + a[2], d02 = a[0] - a[2];
          cplx s13 = a[1] + a[3], d13 = a[1] - a[3];
          cplx jd13 = (sign < 0) ? (-I * d13) : (I * d13);
          dst[k * m + q] = s02 + s13;
          dst[(k + l) * m + q] = d02 + jd13;
          dst[(k + 2 * l) * m + q] = s02 - s13;
          dst[(k + 3 * l) * m + q] = d02 - jd13;
        } else {
          for (int u = 0; u < r; u++) {
            cplx acc = 0;
            for (int t = 0; t < r; t++) {
              cplx w = P->tw[((t * u) % r) * (n / r)];
              if (sign > 0) w = conj(w);
              acc += a[t] * w;
            }
            dst[(k + l * u) * m + q] = acc;
          }
        }
      }
    }
    cplx *tmp = src; src = dst; dst = tmp;
    l = lr;
  }
  if (src != x) memcpy(x, src, sizeof(cplx) * n);
}

void oracle_fft(double *data, int n, int sign) {
  offt_plan *P = offt_create(n);
  cplx *work = (cplx *)malloc(sizeof(cplx) * n);
  offt_exec(P, (cplx *)data, work, sign);
  free(work);
  offt_destroy(P);
}

/* fftw_plan_many_dft(1,&len,howmany,data,NULL,howmany,1,...) as used at
 * reconstruction.c:521-524, 556-559, 607-610: element t of transform b is
 * data[t*howmany + b]. */
void oracle_fft_strided(double *data, int len, int howmany, int sign) {
  offt_plan *P = offt_create(len);
  cplx *d = (cplx *)data;
#pragma omp parallel
  {
    cplx *line = (cplx *)malloc(sizeof(cplx) * len);
    cplx *work = (cplx *)malloc(sizeof(cplx) * len);
#pragma omp for schedule(static)
    for (int b = 0; b < howmany; b++) {
      for (int t = 0; t < len; t++) line[t] = d[(size_t)t * howmany + b];
      offt_exec(P, line, work, sign);
      for (int t = 0; t < len; t++) d[(size_t)t * howmany + b] = line[t];
    }
    free(line);
    free(work);
  }
  offt_destroy(P);
}

/* ------------------------------------------------------------------------ */
/* projection FT and NDFT                                                    */
/* ------------------------------------------------------------------------ */

/* reconstruction.c:306-367 with nfft_trafo replaced by its exact sum:
 * tomo[t][j][i] = sum_{r,c} P_t[r][c] exp(-2 pi i ((c-nx/2) u_i + (r-nz/2) w_j)),
 * (u_i, w_j) = uniform_sample_2D(np, nz). */
void oracle_projection_ft(int nx, int nz, int np, int nmeas, const double *proj, double *tomo) {
  double *x = (double *)malloc(sizeof(double) * 2 * np * nz);
  oracle_uniform_nodes(np, nz, x);
  cplx *ex = (cplx *)malloc(sizeof(cplx) * nx * np); /* [c][i] */
  cplx *ez = (cplx *)malloc(sizeof(cplx) * nz * nz); /* [r][j] */
  for (int c = 0; c < nx; c++)
    for (int i = 0; i < np; i++) {
      double a = -TWO_PI * (double)(c - nx / 2) * x[2 * i + 0];
      ex[c * np + i] = cos(a) + I * sin(a);
    }
  for (int r = 0; r < nz; r++)
    for (int j = 0; j < nz; j++) {
      double a = -TWO_PI * (double)(r - nz / 2) * x[2 * (j * np) + 1];
      ez[r * nz + j] = cos(a) + I * sin(a);
    }
  cplx *T = (cplx *)tomo;
#pragma omp parallel for schedule(static)
  for (int t = 0; t < nmeas; t++) {
    const double *P = proj + (size_t)t * nz * nx;
    cplx *Q = (cplx *)calloc((size_t)nz * np, sizeof(cplx));
    for (int r = 0; r < nz; r++)
      for (int c = 0; c < nx; c++) {
        double v = P[r * nx + c];
        if (v == 0.0) continue;
        for (int i = 0; i < np; i++) Q[r * np + i] += v * ex[c * np + i];
      }
    cplx *Tt = T + (size_t)t * nz * np;
    for (int j = 0; j < nz; j++)
      for (int i = 0; i < np; i++) Tt[j * np + i] = 0;
    for (int r = 0; r < nz; r++)
      for (int j = 0; j < nz; j++) {
        cplx e = ez[r * nz + j];
        for (int i = 0; i < np; i++) Tt[j * np + i] += Q[r * np + i] * e;
      }
    free(Q);
  }
  free(ex);
  free(ez);
  free(x);
}

/* NFFT3's nfft_trafo definition (sign as projection_sim.c:539-546):
 * f[m] = sum_{k0,k1} fhat[k0][k1] exp(-2 pi i ((k0-N/2) x_m0 + (k1-N/2) x_m1)) */
void oracle_ndft_trafo(int N, int M, const double *x, const double *fhat, double *f) {
  const cplx *fh = (const cplx *)fhat;
  cplx *fo = (cplx *)f;
#pragma omp parallel
  {
    cplx *e0 = (cplx *)malloc(sizeof(cplx) * N), *e1 = (cplx *)malloc(sizeof(cplx) * N);
#pragma omp for schedule(static)
    for (int m = 0; m < M; m++) {
      for (int k = 0; k < N; k++) {
        double a0 = -TWO_PI * (double)(k - N / 2) * x[2 * m + 0];
        double a1 = -TWO_PI * (double)(k - N / 2) * x[2 * m + 1];
        e0[k] = cos(a0) + I * sin(a0);
        e1[k] = cos(a1) + I * sin(a1);
      }
      cplx acc = 0;
      for (int k0 = 0; k0 < N; k0++) {
        cplx row = 0;
        for (int k1 = 0; k1 < N; k1++) row += fh[k0 * N + k1] * e1[k1];
        acc += row * e0[k0];
      }
      fo[m] = acc;
    }
    free(e0);
    free(e1);
  }
}

/* nfft_adjoint definition: fhat[k0][k1] = sum_m f[m] exp(+2 pi i (...)) */
void oracle_ndft_adjoint(int N, int M, const double *x, const double *f, double *fhat) {
  const cplx *fi = (const cplx *)f;
  cplx *fh = (cplx *)fhat;
  for (int k = 0; k < N * N; k++) fh[k] = 0;
  cplx *e0 = (cplx *)malloc(sizeof(cplx) * N), *e1 = (cplx *)malloc(sizeof(cplx) * N);
  for (int m = 0; m < M; m++) {
    for (int k = 0; k < N; k++) {
      double a0 = TWO_PI * (double)(k - N / 2) * x[2 * m + 0];
      double a1 = TWO_PI * (double)(k - N / 2) * x[2 * m + 1];
      e0[k] = cos(a0) + I * sin(a0);
      e1[k] = cos(a1) + I * sin(a1);
    }
    for (int k0 = 0; k0 < N; k0++) {
      cplx a = fi[m] * e0[k0];
      for (int k1 = 0; k1 < N; k1++) fh[k0 * N + k1] += a * e1[k1];
    }
  }
  free(e0);
  free(e1);
}

/* ------------------------------------------------------------------------ */
/* Kaiser-Bessel NFFT                                                        */
/* ------------------------------------------------------------------------ */

/* modified Bessel function I0 by its power series (all terms positive) */
static double bessel_i0(double x) {
  double q = 0.25 * x * x, term = 1.0, sum = 1.0;
  for (int k = 1; k < 500; k++) {
    term *= q / ((double)k * (double)k);
    sum += term;
    if (term < 1e-17 * sum) break;
  }
  return sum;
}

/* shape parameter for W taps at oversampling sigma (Beatty, Nishimura and
 * Pauly, IEEE TMI 24(6) 2005, eq. 5) */
double oracle_kb_beta(int W, double sigma) {
  double a = (double)W / sigma * (sigma - 0.5);
  return M_PI * sqrt(a * a - 0.8);
}

/* window psi(d) = I0(beta sqrt(1-(2d/W)^2)) / I0(beta), |d| <= W/2 */
static double kb_psi(double d, int W, double beta, double i0b) {
  double t = 2.0 * d / W;
  double s = 1.0 - t * t;
  if (s < 0) return 0.0;
  return bessel_i0(beta * sqrt(s)) / i0b;
}

/* continuous Fourier transform of psi at frequency nu (cycles per cell) */
static double kb_psihat(double nu, int W, double beta, double i0b) {
  double a = M_PI * W * nu;
  double s = beta * beta - a * a;
  double v;
  if (s > 1e-300) {
    double r = sqrt(s);
    v = sinh(r) / r;
  } else if (s < -1e-300) {
    double r = sqrt(-s);
    v = sin(r) / r;
  } else {
    v = 1.0;
  }
  return W * v / i0b;
}

struct oracle_nfft {
  int N0, N1, n0, n1, W, M;
  double *psi0, *psi1; /* [M][W] */
  int *c0, *c1;        /* first tap cell (unwrapped) */
  double *dec0, *dec1; /* [N0], [N1]: 1/psihat((k-N/2)/n) */
  offt_plan *f0, *f1;  /* column (n0) and row (n1) transforms */
};

typedef struct {
  cplx *g;    /* n0*n1 */
  cplx *line; /* max(n0,n1) */
  cplx *work;
} nfft_ws;

static nfft_ws *nfft_ws_create(const oracle_nfft *p) {
  nfft_ws *w = (nfft_ws *)malloc(sizeof(nfft_ws));
  int L = p->n0 > p->n1 ? p->n0 : p->n1;
  w->g = (cplx *)malloc(sizeof(cplx) * (size_t)p->n0 * p->n1);
  w->line = (cplx *)malloc(sizeof(cplx) * L);
  w->work = (cplx *)malloc(sizeof(cplx) * L);
  return w;
}

static void nfft_ws_destroy(nfft_ws *w) {
  free(w->g);
  free(w->line);
  free(w->work);
  free(w);
}

static void kb_node_taps(double xi, int n, int W, double beta, double i0b, int *c, double *psi) {
  double u = xi * n;
  int fl = (int)floor(u);
  *c = fl - W / 2 + 1;
  for (int t = 0; t < W; t++) psi[t] = kb_psi(u - (double)(*c + t), W, beta, i0b);
}

oracle_nfft *oracle_nfft_create(int N0, int N1, int n0, int n1, int W, int M, const double *x) {
  oracle_nfft *p = (oracle_nfft *)calloc(1, sizeof(oracle_nfft));
  p->N0 = N0; p->N1 = N1; p->n0 = n0; p->n1 = n1; p->W = W; p->M = M;
  double b0 = oracle_kb_beta(W, (double)n0 / N0), b1 = oracle_kb_beta(W, (double)n1 / N1);
  double i0b0 = bessel_i0(b0), i0b1 = bessel_i0(b1);
  p->psi0 = (double *)malloc(sizeof(double) * (size_t)M * W);
  p->psi1 = (double *)malloc(sizeof(double) * (size_t)M * W);
  p->c0 = (int *)malloc(sizeof(int) * M);
  p->c1 = (int *)malloc(sizeof(int) * M);
  for (int m = 0; m < M; m++) {
    kb_node_taps(x[2 * m + 0], n0, W, b0, i0b0, &p->c0[m], &p->psi0[(size_t)m * W]);
    kb_node_taps(x[2 * m + 1], n1, W, b1, i0b1, &p->c1[m], &p->psi1[(size_t)m * W]);
  }
  p->dec0 = (double *)malloc(sizeof(double) * N0);
  p->dec1 = (double *)malloc(sizeof(double) * N1);
  for (int k = 0; k < N0; k++) p->dec0[k] = 1.0 / kb_psihat((double)(k - N0 / 2) / n0, W, b0, i0b0);
  for (int k = 0; k < N1; k++) p->dec1[k] = 1.0 / kb_psihat((double)(k - N1 / 2) / n1, W, b1, i0b1);
  p->f0 = offt_create(n0);
  p->f1 = offt_create(n1);
  return p;
}

void oracle_nfft_destroy(oracle_nfft *p) {
  if (!p) return;
  free(p->psi0); free(p->psi1); free(p->c0); free(p->c1);
  free(p->dec0); free(p->dec1);
  offt_destroy(p->f0); offt_destroy(p->f1);
  free(p);
}

static inline int pmod(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}

/* forward: deconvolve, zero-pad, 2D FFT (e^-), W x W window interpolation */
static void nfft_trafo_ws(const oracle_nfft *p, nfft_ws *ws, const cplx *fhat, cplx *f) {
  int N0 = p->N0, N1 = p->N1, n0 = p->n0, n1 = p->n1, W = p->W;
  cplx *g = ws->g;
  memset(g, 0, sizeof(cplx) * (size_t)n0 * n1);
  for (int k0 = 0; k0 < N0; k0++) {
    cplx *row = g + (size_t)pmod(k0 - N0 / 2, n0) * n1;
    for (int k1 = 0; k1 < N1; k1++)
      row[pmod(k1 - N1 / 2, n1)] = fhat[k0 * N1 + k1] * (p->dec0[k0] * p->dec1[k1]);
    offt_exec(p->f1, row, ws->work, -1);
  }
  for (int c = 0; c < n1; c++) {
    for (int r = 0; r < n0; r++) ws->line[r] = g[(size_t)r * n1 + c];
    offt_exec(p->f0, ws->line, ws->work, -1);
    for (int r = 0; r < n0; r++) g[(size_t)r * n1 + c] = ws->line[r];
  }
  for (int m = 0; m < p->M; m++) {
    const double *a0 = p->psi0 + (size_t)m * W, *a1 = p->psi1 + (size_t)m * W;
    cplx acc = 0;
    for (int t0 = 0; t0 < W; t0++) {
      const cplx *row = g + (size_t)pmod(p->c0[m] + t0, n0) * n1;
      cplx racc = 0;
      for (int t1 = 0; t1 < W; t1++) racc += a1[t1] * row[pmod(p->c1[m] + t1, n1)];
      acc += a0[t0] * racc;
    }
    f[m] = acc;
  }
}

/* adjoint: spread, 2D FFT (e^+), crop, deconvolve */
static void nfft_adjoint_ws(const oracle_nfft *p, nfft_ws *ws, const cplx *f, cplx *fhat) {
  int N0 = p->N0, N1 = p->N1, n0 = p->n0, n1 = p->n1, W = p->W;
  cplx *g = ws->g;
  memset(g, 0, sizeof(cplx) * (size_t)n0 * n1);
  for (int m = 0; m < p->M; m++) {
    const double *a0 = p->psi0 + (size_t)m * W, *a1 = p->psi1 + (size_t)m * W;
    for (int t0 = 0; t0 < W; t0++) {
      cplx *row = g + (size_t)pmod(p->c0[m] + t0, n0) * n1;
      cplx v = a0[t0] * f[m];
      for (int t1 = 0; t1 < W; t1++) row[pmod(p->c1[m] + t1, n1)] += a1[t1] * v;
    }
  }
  for (int c = 0; c < n1; c++) {
    for (int r = 0; r < n0; r++) ws->line[r] = g[(size_t)r * n1 + c];
    offt_exec(p->f0, ws->line, ws->work, +1);
    for (int r = 0; r < n0; r++) g[(size_t)r * n1 + c] = ws->line[r];
  }
  for (int k0 = 0; k0 < N0; k0++) {
    cplx *row = g + (size_t)pmod(k0 - N0 / 2, n0) * n1;
    offt_exec(p->f1, row, ws->work, +1);
    for (int k1 = 0; k1 < N1; k1++)
      fhat[k0 * N1 + k1] = row[pmod(k1 - N1 / 2, n1)] * (p->dec0[k0] * p->dec1[k1]);
  }
}

void oracle_nfft_trafo(oracle_nfft *p, const double *fhat, double *f) {
  nfft_ws *ws = nfft_ws_create(p);
  nfft_trafo_ws(p, ws, (const cplx *)fhat, (cplx *)f);
  nfft_ws_destroy(ws);
}

void oracle_nfft_adjoint(oracle_nfft *p, const double *f, double *fhat) {
  nfft_ws *ws = nfft_ws_create(p);
  nfft_adjoint_ws(p, ws, (const cplx *)f, (cplx *)fhat);
  nfft_ws_destroy(ws);
}

/* ------------------------------------------------------------------------ */
/* CGNR, NFFT3 solver order (reconstruction.c:413, 506-510, 592-595)          */
/* ------------------------------------------------------------------------ */

typedef struct {
  cplx *r, *v, *tmp; /* M */
  cplx *z, *pp;      /* N */
} cg_ws;

static void cgnr_ws(const oracle_nfft *p, nfft_ws *ws, cg_ws *cw, const double *w, const cplx *y,
                    cplx *fh, int nsteps) {
  int M = p->M, Nt = p->N0 * p->N1;
  cplx *r = cw->r, *v = cw->v, *tmp = cw->tmp, *z = cw->z, *pp = cw->pp;
  /* before_loop: r = y - A f_hat_iter; z = A^H (W r); p = z */
  nfft_trafo_ws(p, ws, fh, v);
  for (int k = 0; k < M; k++) r[k] = y[k] - v[k];
  for (int k = 0; k < M; k++) tmp[k] = w ? w[k] * r[k] : r[k];
  nfft_adjoint_ws(p, ws, tmp, z);
  double dot_z = 0;
  for (int k = 0; k < Nt; k++) dot_z += creal(z[k]) * creal(z[k]) + cimag(z[k]) * cimag(z[k]);
  for (int k = 0; k < Nt; k++) pp[k] = z[k];
  for (int it = 0; it < nsteps; it++) {
    nfft_trafo_ws(p, ws, pp, v);
    double dot_v = 0;
    for (int k = 0; k < M; k++) {
      double a = creal(v[k]) * creal(v[k]) + cimag(v[k]) * cimag(v[k]);
      dot_v += w ? w[k] * a : a;
    }
    double alpha = dot_z / dot_v;
    for (int k = 0; k < Nt; k++) fh[k] += alpha * pp[k];
    for (int k = 0; k < M; k++) r[k] += (-alpha) * v[k];
    for (int k = 0; k < M; k++) tmp[k] = w ? w[k] * r[k] : r[k];
    nfft_adjoint_ws(p, ws, tmp, z);
    double dot_z_old = dot_z;
    dot_z = 0;
    for (int k = 0; k < Nt; k++) dot_z += creal(z[k]) * creal(z[k]) + cimag(z[k]) * cimag(z[k]);
    double beta = dot_z / dot_z_old;
    for (int k = 0; k < Nt; k++) pp[k] = beta * pp[k] + z[k];
  }
}

static cg_ws *cg_ws_create(int M, int Nt) {
  cg_ws *c = (cg_ws *)malloc(sizeof(cg_ws));
  c->r = (cplx *)malloc(sizeof(cplx) * M);
  c->v = (cplx *)malloc(sizeof(cplx) * M);
  c->tmp = (cplx *)malloc(sizeof(cplx) * M);
  c->z = (cplx *)malloc(sizeof(cplx) * Nt);
  c->pp = (cplx *)malloc(sizeof(cplx) * Nt);
  return c;
}

static void cg_ws_destroy(cg_ws *c) {
  free(c->r); free(c->v); free(c->tmp); free(c->z); free(c->pp);
  free(c);
}

void oracle_cgnr(oracle_nfft *p, const double *w, const double *y, double *fhat, int nsteps) {
  nfft_ws *ws = nfft_ws_create(p);
  cg_ws *cw = cg_ws_create(p->M, p->N0 * p->N1);
  cgnr_ws(p, ws, cw, w, (const cplx *)y, (cplx *)fhat, nsteps);
  cg_ws_destroy(cw);
  nfft_ws_destroy(ws);
}

/* ------------------------------------------------------------------------ */
/* denoising (denoise.c)                                                     */
/* ------------------------------------------------------------------------ */

#define IDX(z, y, x) (((size_t)(z) * ny + (y)) * nx + (x))

/* denoise.c:97-121 (and :185-209): the per-voxel gradient of the discrete TV,
 * for rows ylo..yhi inclusive.  The reference's (i,j,k) are (z,y,x) here. */
void oracle_tv_grad(const double *u, int nz, int ny, int nx, int ylo, int yhi, double *grad) {
  const double eps = 1e-8;
#pragma omp parallel for schedule(static)
  for (int z = 0; z < nz; z++) {
    int zp = (z == nz - 1) ? nz - 1 : z + 1, zm = (z == 0) ? 0 : z - 1;
    for (int y = 0; y < ny; y++) {
      if (y < ylo || y > yhi) continue;
      int yp = (y == ny - 1) ? ny - 1 : y + 1, ym = (y == 0) ? 0 : y - 1;
      for (int x = 0; x < nx; x++) {
        int xp = (x == nx - 1) ? nx - 1 : x + 1, xm = (x == 0) ? 0 : x - 1;
        double c = u[IDX(z, y, x)];
        double t1 = c - u[IDX(zm, y, x)];
        double t2 = c - u[IDX(z, ym, x)];
        double t3 = c - u[IDX(z, y, xm)];
        double t4 = u[IDX(zp, y, x)] - c;
        double t5 = u[IDX(zp, y, x)] - u[IDX(zp, ym, x)];
        double t6 = u[IDX(zp, y, x)] - u[IDX(zp, y, xm)];
        double t7 = u[IDX(z, yp, x)] - u[IDX(zm, yp, x)];
        double t8 = u[IDX(z, yp, x)] - c;
        double t9 = u[IDX(z, yp, x)] - u[IDX(z, yp, xm)];
        double t10 = u[IDX(z, y, xp)] - u[IDX(zm, y, xp)];
        double t11 = u[IDX(z, y, xp)] - u[IDX(z, ym, xp)];
        double t12 = u[IDX(z, y, xp)] - c;
        grad[IDX(z, y, x)] = (t1 + t2 + t3) / sqrt(eps + t1 * t1 + t2 * t2 + t3 * t3) -
                             t4 / sqrt(eps + t4 * t4 + t5 * t5 + t6 * t6) -
                             t8 / sqrt(eps + t7 * t7 + t8 * t8 + t9 * t9) -
                             t12 / sqrt(eps + t10 * t10 + t11 * t11 + t12 * t12);
      }
    }
  }
}

/* denoise.c:67-151 / 153-249: niter sweeps of
 *   g = grad / sqrt(sum grad^2);  dA = sum_{u<0} u^2 (pre-update u);
 *   u -= lam * sqrt(dA) * g  over rows ylo..yhi. */
static void tv_sweeps(double *u, float lam, int niter, int nz, int ny, int nx, int ylo, int yhi) {
  size_t V = (size_t)nz * ny * nx;
  double *grad = (double *)malloc(sizeof(double) * V);
  double *part = (double *)malloc(sizeof(double) * 2 * nz);
  int y0 = ylo < 0 ? 0 : ylo, y1 = yhi > ny - 1 ? ny - 1 : yhi;
  while (niter--) {
    oracle_tv_grad(u, nz, ny, nx, ylo, yhi, grad);
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; z++) {
      double s = 0;
      for (int y = y0; y <= y1; y++)
        for (int x = 0; x < nx; x++) s += grad[IDX(z, y, x)] * grad[IDX(z, y, x)];
      part[z] = s;
    }
    double tmp0 = 0;
    for (int z = 0; z < nz; z++) tmp0 += part[z];
    double s1 = sqrt(tmp0);
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; z++) {
      double d = 0;
      for (int y = y0; y <= y1; y++)
        for (int x = 0; x < nx; x++) {
          grad[IDX(z, y, x)] /= s1;
          double v = u[IDX(z, y, x)];
          if (v < 0) d += v * v;
        }
      part[nz + z] = d;
    }
    double dA = 0;
    for (int z = 0; z < nz; z++) dA += part[nz + z];
    double s2 = sqrt(dA);
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; z++)
      for (int y = y0; y <= y1; y++)
        for (int x = 0; x < nx; x++) u[IDX(z, y, x)] -= lam * s2 * grad[IDX(z, y, x)];
  }
  free(part);
  free(grad);
}

void oracle_tv(double *u, float lam, int niter, int nz, int ny, int nx) {
  tv_sweeps(u, lam, niter, nz, ny, nx, 0, ny - 1);
}

/* band test of denoise.c:180 with C integer division */
void oracle_tv_height(double *u, float lam, int niter, int nz, int ny, int nx, int nheight,
                      int ncenter) {
  tv_sweeps(u, lam, niter, nz, ny, nx, (2 * ncenter - nheight) / 2, (2 * ncenter + nheight) / 2);
}

/* denoise.c:320-378: per z plane, in-place Gauss-Seidel in (y, x) order.
 * The reference clamps x with the z extent (:349-350); that is only in
 * bounds when nz == nx, so other shapes are rejected. */
int oracle_diffusion2d(double *u, float lambda, float kappa, int niter, int nz, int ny, int nx) {
  if (nz != nx) return -1;
  double Ksq = pow(kappa, 2);
  while (niter--) {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; k++) {
      for (int j = 0; j < ny; j++) {
        int jp = (j == ny - 1) ? ny - 1 : j + 1, jm = (j == 0) ? 0 : j - 1;
        for (int i = 0; i < nx; i++) {
          int ip = (i == nz - 1) ? nz - 1 : i + 1, im = (i == 0) ? 0 : i - 1;
          double c = u[IDX(k, j, i)];
          double gN = u[IDX(k, jm, i)] - c;
          double gS = u[IDX(k, jp, i)] - c;
          double gE = u[IDX(k, j, ip)] - c;
          double gW = u[IDX(k, j, im)] - c;
          double cN = 1.0 / (1.0 + pow(gN, 2) / Ksq);
          double cS = 1.0 / (1.0 + pow(gS, 2) / Ksq);
          double cE = 1.0 / (1.0 + pow(gE, 2) / Ksq);
          double cW = 1.0 / (1.0 + pow(gW, 2) / Ksq);
          u[IDX(k, j, i)] += lambda * (cN * gN + cS * gS + cE * gE + cW * gW);
        }
      }
    }
  }
  return 0;
}
#undef IDX

/* reconstruction.c:672-700 */
void oracle_constraint(double *img, const double *img_last, int nz, int np, int nx, int is_hio,
                       int config_height, int nheight, int ncenter) {
  size_t plane = (size_t)np * np;
  for (int j = 0; j < nz; j++)
    for (int i = 0; i < np * np; i++) {
      int yi = i / np, xi = i % np;
      size_t o = j * plane + i;
      int inside;
      if (config_height)
        inside = (yi >= (2 * ncenter - nheight) / 2 && yi <= (2 * ncenter + nheight) / 2);
      else
        inside = (yi >= (np - nx) / 2) && (yi < (np + nx) / 2) && (xi >= (np - nx) / 2) &&
                 (xi < (np + nx) / 2);
      if (inside) {
        if (img[o] < 0) img[o] = is_hio ? img_last[o] - 0.9 * img[o] : 0;
      } else {
        img[o] = is_hio ? img_last[o] - 0.9 * img[o] : 0;
      }
    }
}

/* ------------------------------------------------------------------------ */
/* end-to-end                                                                */
/* ------------------------------------------------------------------------ */

static int next_pow2(int n) {
  int p = 1;
  while (p < n) p <<= 1;
  return p;
}

typedef struct {
  int nmeas, nang, M;
  double pad_min, pad_max;
  double *x, *w;
  oracle_nfft *plan;
} geo_t;

static int geo_init(const oracle_params *p, geo_t *g) {
  if (p->np <= 0 || p->nz <= 0 || (p->np & 1) || (p->nz & 1)) return -1;
  g->nmeas = (int)((p->gam_max - p->gam_min) / p->interval + 0.5) + 1;
  oracle_pad_range(p->gam_min, p->interval, &g->pad_min, &g->pad_max);
  g->M = oracle_polar_nodes(g->pad_min, g->pad_max, p->padangle, p->np, NULL);
  g->nang = g->M / p->np;
  g->x = (double *)malloc(sizeof(double) * 2 * g->M);
  oracle_polar_nodes(g->pad_min, g->pad_max, p->padangle, p->np, g->x);
  g->w = NULL;
  if (p->weight) {
    g->w = (double *)malloc(sizeof(double) * g->M);
    oracle_weights(p->np, g->nang, p->padangle, g->w);
  }
  int n = 2 * next_pow2(p->np);
  g->plan = oracle_nfft_create(p->np, p->np, n, n, 14, g->M, g->x);
  return 0;
}

static void geo_free(geo_t *g) {
  free(g->x);
  free(g->w);
  oracle_nfft_destroy(g->plan);
}

int oracle_assemble(const oracle_params *p, const double *proj, double *slices, int *kind) {
  geo_t g;
  if (geo_init(p, &g)) return -1;
  int np = p->np, nz = p->nz, M = g.M;
  double *tomo = (double *)malloc(sizeof(double) * 2 * (size_t)g.nmeas * nz * np);
  oracle_projection_ft(p->nx, nz, np, g.nmeas, proj, tomo);
  int *order = (int *)malloc(sizeof(int) * M);
  oracle_mask(np, g.nang, g.pad_min, p->padangle, p->gam_min, p->gam_max, p->interval, kind, order);
  cplx *s = (cplx *)slices;
  const cplx *t = (const cplx *)tomo;
  for (int j = 0; j < nz; j++)
    for (int i = 0; i < M; i++)
      s[(size_t)j * M + i] = (kind[i] == 1) ? t[(size_t)order[i] * nz * np + j * np + i % np] : 0;
  free(order);
  free(tomo);
  geo_free(&g);
  return 0;
}

/* reconstruction.c:489-539 */
int oracle_initial(const oracle_params *p, const double *slices, double *img0) {
  geo_t g;
  if (geo_init(p, &g)) return -1;
  int np = p->np, nz = p->nz, M = g.M, Nt = np * np;
  size_t V = (size_t)Nt * nz;
  cplx *data = (cplx *)malloc(sizeof(cplx) * V);
  const cplx *s = (const cplx *)slices;
#pragma omp parallel
  {
    nfft_ws *ws = nfft_ws_create(g.plan);
    cg_ws *cw = cg_ws_create(M, Nt);
    cplx *fh = (cplx *)malloc(sizeof(cplx) * Nt);
#pragma omp for schedule(dynamic, 1)
    for (int j = 0; j < nz; j++) {
      for (int i = 0; i < Nt; i++) fh[i] = 0.0;
      cgnr_ws(g.plan, ws, cw, g.w, s + (size_t)j * M, fh, 10);
      for (int i = 0; i < Nt; i++) data[((size_t)j * Nt + i + V / 2) % V] = fh[i];
    }
    free(fh);
    cg_ws_destroy(cw);
    nfft_ws_destroy(ws);
  }
  oracle_fft_strided((double *)data, nz, Nt, +1);
  for (int j = 0; j < nz; j++)
    for (int i = 0; i < Nt; i++) img0[(size_t)((j + nz / 2) % nz) * Nt + i] = creal(data[(size_t)j * Nt + i]) / nz;
  free(data);
  geo_free(&g);
  return 0;
}

/* per-plane Fourier step of reconstruction.c:565-601 for planes j0 <= j < j1:
 * data (z-FFT'd volume, [nz][np*np]) in/out, img the current real volume,
 * d1/d2 per-plane R-factor sums */
static void fourier_planes(const oracle_params *p, const geo_t *g, const double *img, double complex *data,
                           const double complex *s, const int *kind, int j0, int j1, double *d1,
                           double *d2) {
  int np = p->np, nz = p->nz, M = g->M, Nt = np * np;
#pragma omp parallel
  {
    nfft_ws *ws = nfft_ws_create(g->plan);
    cg_ws *cw = cg_ws_create(M, Nt);
    cplx *fh = (cplx *)malloc(sizeof(cplx) * Nt);
    cplx *f = (cplx *)malloc(sizeof(cplx) * M);
    cplx *y = (cplx *)malloc(sizeof(cplx) * M);
#pragma omp for schedule(dynamic, 1)
    for (int j = j0; j < j1; j++) {
      size_t pl = (size_t)((j + nz / 2) % nz) * Nt;
      for (int i = 0; i < Nt; i++) fh[i] = data[pl + i];
      nfft_trafo_ws(g->plan, ws, fh, f);
      double a1 = 0, a2 = 0;
      for (int i = 0; i < M; i++) {
        cplx sv = s[(size_t)j * M + i];
        if (kind[i] == 1) {
          y[i] = sv;
          a1 += pow(creal(f[i] - sv), 2) + pow(cimag(f[i] - sv), 2);
          a2 += pow(creal(f[i] + sv), 2) + pow(cimag(f[i] + sv), 2);
        } else {
          y[i] = f[i];
        }
      }
      d1[j] = a1;
      d2[j] = a2;
      for (int i = 0; i < Nt; i++) fh[i] = img[pl + i];
      cgnr_ws(g->plan, ws, cw, g->w, y, fh, 10);
      for (int i = 0; i < Nt; i++) data[pl + i] = fh[i];
    }
    free(fh); free(f); free(y);
    cg_ws_destroy(cw);
    nfft_ws_destroy(ws);
  }
}

/* reconstruction.c:545-700, one pass of the loop body with counter nn */
int oracle_iterate(const oracle_params *p, double *img, double *img_last_out, const double *slices,
                   const int *kind, int nn, double *rfactor) {
  geo_t g;
  if (geo_init(p, &g)) return -1;
  int np = p->np, nz = p->nz, Nt = np * np;
  size_t V = (size_t)Nt * nz;
  cplx *data = (cplx *)malloc(sizeof(cplx) * V);
  double *img_last = (double *)malloc(sizeof(double) * V);
  double *d1 = (double *)calloc(nz, sizeof(double)), *d2 = (double *)calloc(nz, sizeof(double));
  for (int j = 0; j < nz; j++)
    for (int i = 0; i < Nt; i++) {
      data[(size_t)j * Nt + i] = img[(size_t)((j + nz / 2) % nz) * Nt + i];
      img_last[(size_t)j * Nt + i] = img[(size_t)j * Nt + i];
    }
  oracle_fft_strided((double *)data, nz, Nt, -1);
  fourier_planes(p, &g, img, data, (const cplx *)slices, kind, 0, nz, d1, d2);
  double diff1 = 0, diff2 = 0;
  for (int j = 0; j < nz; j++) { diff1 += d1[j]; diff2 += d2[j]; }
  if (rfactor) *rfactor = sqrt(diff1 / diff2);
  oracle_fft_strided((double *)data, nz, Nt, +1);
  for (int j = 0; j < nz; j++)
    for (int i = 0; i < Nt; i++)
      img[(size_t)j * Nt + i] = creal(data[(size_t)((j + nz / 2) % nz) * Nt + i]) / nz;
  int rc = 0;
  if (p->denoise) {
    if (p->method == 0) {
      rc = oracle_diffusion2d(img, p->diffusion_lambda > 0 ? p->diffusion_lambda : 0.25f, 0.005f, 80, nz, np, np);
    } else if (p->config_height && !p->tv_full) {
      oracle_tv_height(img, 0.2f, 40, nz, np, np, p->nheight, p->ncenter);
    } else {
      oracle_tv(img, 0.2f, 40, nz, np, np);
    }
  }
  oracle_constraint(img, img_last, nz, np, p->nx, nn < p->nhio, p->config_height, p->nheight,
                    p->ncenter);
  if (img_last_out) memcpy(img_last_out, img_last, sizeof(double) * V);
  free(d1); free(d2);
  free(img_last);
  free(data);
  geo_free(&g);
  return rc;
}

/* timing sample of one pass: the Fourier step of planes [j0, j1) (plan set-up
 * excluded); returns seconds */
double oracle_time_fourier_planes(const oracle_params *p, const double *img, const double *slices,
                                  const int *kind, int j0, int j1) {
  geo_t g;
  if (geo_init(p, &g)) return -1;
  int nz = p->nz, Nt = p->np * p->np;
  size_t V = (size_t)Nt * nz;
  cplx *data = (cplx *)malloc(sizeof(cplx) * V);
  for (size_t i = 0; i < V; i++) data[i] = img[i];
  double *d1 = (double *)calloc(nz, sizeof(double)), *d2 = (double *)calloc(nz, sizeof(double));
  double t0 = oracle_wtime();
  fourier_planes(p, &g, img, data, (const cplx *)slices, kind, j0, j1, d1, d2);
  double t = oracle_wtime() - t0;
  free(d1); free(d2); free(data);
  geo_free(&g);
  return t;
}

int oracle_reconstruct(const oracle_params *p, const double *proj, double *img0, double *img,
                       double *rtrace) {
  geo_t g;
  if (geo_init(p, &g)) return -1;
  int M = g.M;
  geo_free(&g);
  if (p->nthreads > 0) oracle_set_threads(p->nthreads);
  size_t V = (size_t)p->np * p->np * p->nz;
  double *slices = (double *)malloc(sizeof(double) * 2 * (size_t)p->nz * M);
  int *kind = (int *)malloc(sizeof(int) * M);
  double *work = img ? img : (double *)malloc(sizeof(double) * V);
  int rc = oracle_assemble(p, proj, slices, kind);
  if (!rc) rc = oracle_initial(p, slices, work);
  if (!rc && img0) memcpy(img0, work, sizeof(double) * V);
  for (int nn = 0; !rc && nn < p->nhio + p->ner; nn++) {
    double r = 0;
    rc = oracle_iterate(p, work, NULL, slices, kind, nn, &r);
    if (rtrace) rtrace[nn] = r;
  }
  if (work != img) free(work);
  free(kind);
  free(slices);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* simulator                                                                 */
/* ------------------------------------------------------------------------ */

/* projection_sim.c:282-405 (-m 2D1D), without noise */
int oracle_project_2d1d(int nx, int ny, int nz, int np, double gam_min, double gam_max,
                        double interval, const double *density, double *proj) {
  if ((nx & 1) || (ny & 1) || (nz & 1) || (np & 1)) return -1;
  int nslices = (int)((gam_max - gam_min) / interval + 0.5) + 1;
  size_t V = (size_t)nx * ny * nz;
  cplx *data = (cplx *)malloc(sizeof(cplx) * V);
  for (size_t i = 0; i < V; i++) data[i] = density[i];
  oracle_fft_strided((double *)data, nz, nx * ny, -1);
  int M = oracle_polar_nodes(gam_min, gam_max, interval, np, NULL);
  double *x = (double *)malloc(sizeof(double) * 2 * M);
  oracle_polar_nodes(gam_min, gam_max, interval, np, x);
  oracle_nfft *p = oracle_nfft_create(ny, nx, 2 * next_pow2(ny), 2 * next_pow2(nx), 14, M, x);
  cplx *slices = (cplx *)malloc(sizeof(cplx) * (size_t)np * nz * nslices);
#pragma omp parallel
  {
    nfft_ws *ws = nfft_ws_create(p);
    cplx *fh = (cplx *)malloc(sizeof(cplx) * nx * ny);
    cplx *f = (cplx *)malloc(sizeof(cplx) * M);
#pragma omp for schedule(dynamic, 1)
    for (int j = 0; j < nz; j++) {
      for (int i = 0; i < nx * ny; i++) fh[i] = data[((size_t)j * nx * ny + i + V / 2) % V];
      nfft_trafo_ws(p, ws, fh, f);
      for (int i = 0; i < M; i++) slices[(size_t)(i / np) * np * nz + j * np + i % np] = f[i];
    }
    free(fh); free(f);
    nfft_ws_destroy(ws);
  }
  oracle_nfft_destroy(p);
  offt_plan *Pr = offt_create(np), *Pc = offt_create(nz);
  cplx *slice = (cplx *)malloc(sizeof(cplx) * np * nz);
  cplx *line = (cplx *)malloc(sizeof(cplx) * (np > nz ? np : nz));
  cplx *work = (cplx *)malloc(sizeof(cplx) * (np > nz ? np : nz));
  for (int j = 0; j < nslices; j++) {
    for (int i = 0; i < np * nz; i++) {
      int row = (i / np + nz / 2) % nz, col = (i % np + np / 2) % np;
      slice[row * np + col] = slices[(size_t)j * np * nz + i];
    }
    for (int r = 0; r < nz; r++) offt_exec(Pr, slice + r * np, work, +1);
    for (int c = 0; c < np; c++) {
      for (int r = 0; r < nz; r++) line[r] = slice[r * np + c];
      offt_exec(Pc, line, work, +1);
      for (int r = 0; r < nz; r++) slice[r * np + c] = line[r];
    }
    for (int i = 0; i < np * nz; i++) {
      int row = i / np, col = i % np, cs = (col + np / 2) % np;
      if (col >= np / 2) continue;
      cplx t = slice[row * np + col];
      slice[row * np + col] = slice[row * np + cs];
      slice[row * np + cs] = t;
    }
    for (int i = 0; i < np * nz; i++) proj[(size_t)j * np * nz + i] = creal(slice[i]) / (np * nz);
  }
  free(slice); free(line); free(work);
  offt_destroy(Pr); offt_destroy(Pc);
  free(slices);
  free(x);
  free(data);
  return 0;
}

/* projection_sim.c:516-601 (-m DFT): brute-force sum with PI = 3.1415926 in
 * the exponent, nodes from equal_angle_sample_3D (grid.c:5-57) */
int oracle_project_dft(int nx, int ny, int nz, int np, double gam_min, double gam_max,
                       double interval, const double *density, double *proj) {
  if ((nz & 1) || (np & 1)) return -1;
  int nslices = (int)((gam_max - gam_min) / interval + 0.5) + 1;
  size_t V = (size_t)nx * ny * nz;
  cplx *slice = (cplx *)malloc(sizeof(cplx) * np * nz);
  offt_plan *Pr = offt_create(np), *Pc = offt_create(nz);
  cplx *line = (cplx *)malloc(sizeof(cplx) * (np > nz ? np : nz));
  cplx *work = (cplx *)malloc(sizeof(cplx) * (np > nz ? np : nz));
  double angle = gam_min - interval;
  for (int k = 0; k < nslices; k++) {
    angle += interval;
    /* grid_slice_3D(np, nz, angle) */
    double *x3 = (double *)malloc(sizeof(double) * 3 * np * nz);
    double ivx = 1.0 / np, ivz = 1.0 / nz, z0 = -0.5 - ivz;
    for (int j = 0; j < nz; j++) {
      z0 += ivz;
      double x0 = -0.5 - ivx;
      for (int i = 0; i < np; i++) {
        x0 += ivx;
        x3[3 * (j * np + i) + 0] = x0 * cos(angle / 180 * REF_PI);
        x3[3 * (j * np + i) + 1] = x0 * sin(angle / 180 * REF_PI);
        x3[3 * (j * np + i) + 2] = z0;
      }
    }
#pragma omp parallel for schedule(static)
    for (int j = 0; j < np * nz; j++) {
      const double *xx = x3 + 3 * j;
      cplx acc = 0;
      for (size_t i = 0; i < V; i++) {
        int col = (int)((i % ((size_t)nx * ny)) % nx) - nx / 2;
        int row = (int)((i % ((size_t)nx * ny)) / nx) - ny / 2;
        int slc = (int)(i / ((size_t)nx * ny)) - nz / 2;
        double ph = 2 * REF_PI * (col * xx[0] + row * xx[1] + slc * xx[2]);
        acc += density[i] * (cos(ph) - sin(ph) * I);
      }
      slice[j] = acc;
    }
    free(x3);
    for (int pass = 0; pass < 2; pass++) {
      if (pass == 1) {
        for (int r = 0; r < nz; r++) offt_exec(Pr, slice + r * np, work, +1);
        for (int c = 0; c < np; c++) {
          for (int r = 0; r < nz; r++) line[r] = slice[r * np + c];
          offt_exec(Pc, line, work, +1);
          for (int r = 0; r < nz; r++) slice[r * np + c] = line[r];
        }
      }
      for (int i = 0; i < np * nz / 2; i++) {
        int row = i / np, col = i % np, rs = (row + nz / 2) % nz, cs = (col + np / 2) % np;
        cplx t = slice[row * np + col];
        slice[row * np + col] = slice[rs * np + cs];
        slice[rs * np + cs] = t;
      }
    }
    for (int i = 0; i < np * nz; i++) proj[(size_t)k * np * nz + i] = creal(slice[i]) / (np * nz);
  }
  free(slice); free(line); free(work);
  offt_destroy(Pr); offt_destroy(Pc);
  return 0;
}
