// engine.hpp -- device-side operators of one reconstruction job.
#pragma once

#include <memory>
#include <vector>

#include "common.hpp"
#include "nudim.h"
#include "fft.hpp"
#include "geom.hpp"
#include "kernels.hpp"

namespace nudim {

// fixed-length register FFT (fft2.hip); false when n has no specialisation
template <typename T>
bool fft2_launch(int n, const cx<T> *in, cx<T> *out, int nlines, const LineIO &li, const LineIO &lo, int sign, int P,
                 hipStream_t st, const cx<T> *tw, const cx<T> *mid, long long mid_so);
// process-wide switch for the fixed-length kernels (on by default)
extern bool g_use_fft2;
// Toeplitz passes with the CGNR vector work (fft2.hip) for n = 2N in
// toep_supported(n); kind is TK_ROWF, TK_COL or TK_ROWI.  P planes from the
// given (offset) pointers; pstride = row stride of cg.part.  pairs: two lanes
// per line in the 48-point stage of n = 1536 (same results to rounding).
bool toep_supported(int n);
template <typename T>
void toep_launch(int kind, int n, int nlines, const cx<T> *in, cx<T> *out, long long in_so, long long in_sk,
                 long long out_so, long long out_sk, int P, hipStream_t st, const cx<T> *tw, const T *mid,
                 const ToepCg<T> &cg, int pstride, bool pairs);

// z-axis transforms of [line][z] volumes on the fixed-length kernels
// (k_zfft2, fft2.hpp); false when n has no specialisation.  The inverse
// adds sum_{s < nk} alpha[s n K + plane] pv[s pstride + i] to its input.
// With K > 1 the real-space side is K volumes vimg elements apart
// (zimg_base) and the Fourier side has n K planes.
template <typename T>
bool zfft2_launch(int n, bool inv, long long nlines, const T *img_in, const cx<T> *g_in, const cx<T> *pv,
                  const double *alpha, cx<T> *fh, cx<T> *f0, T *img_out, const cx<T> *tw, hipStream_t st,
                  int nk = 0, long long pstride = 0, int K = 1, long long vimg = 0);

// FP32 NFFT sizing for a node set (tune.hip): the fastest measured (n, W)
// whose trafo and adjoint stay within 2e-6 relative L2 of the FP64 NFFT at
// the reference's sizes; n <= nice_fft_size(2N).  err / ms < 0: not measured
// or above the tolerance.  trace (may be null) gets every candidate tried.
struct NfftTuning {
  int n, W;
  double err, ms;
};
NfftTuning tune_nfft_fp32(int N, const double *x, int M, hipStream_t st, std::vector<NfftTuning> *trace);

// choose log2 of the plane chunk G of a line transform of length n over P planes
int line_chunk_log2(int n, int P, int elem_bytes, int ept);

template <typename T>
struct LineFft {
  FftDesc fd{};
  DevBuf tw;  // cx<T>[n]
  int n = 0;
  void init(int len);
  // one batched line transform (see k_fft_lines)
  void run(const cx<T> *in, cx<T> *out, int nlines, const LineIO &li, const LineIO &lo, int sign, int P,
           hipStream_t st, const cx<T> *mid = nullptr, long long mid_so = 0) const;
};

// The node tables of one NFFT geometry (N, n, W and the nodes), which every
// operator and job with that geometry shares: window and deconvolution
// tables, line-FFT twiddles and the gather maps.  nfft_geometry() builds them
// once per device and geometry and keeps the most recent ones.
template <typename T>
struct NfftGeom {
  int N = 0, n = 0, W = 0, M = 0;
  LineFft<T> fft;
  DevBuf c0, c1, psi0, psi1, dec;
  DevBuf cell_ptr, ent;  // per-cell gather map (any plane count)
  // the nodes in the order the interpolation visits them: by 8 x 8-cell
  // tile of their first tap, tiles in Morton (Z) order, so that the nodes of
  // a block, and the blocks of an XCD, read the same grid cells from L2
  DevBuf perm;
  // patch lists of the patch-form kernels (spread.hpp; used when P % 64 == 0):
  // the gather's 8 x 4-cell patches (entry-list segments dealt to the XCDs by
  // angular sector, longest first in each, with the split patches' (patch,
  // first slot, count) triples; PatchRec<T, 8, 4> records) and the interpolation's 8 x 16-cell block patches (segments of
  // the nonempty ones, four PatchRec<T> quarter records per entry, each
  // node's entries)
  bool has_patches = false, has_blocks = false;
  int npatch = 0, ngseg = 0, ncomb = 0, nslot = 0, niseg = 0, nbentry = 0;
  DevBuf gseg, gcomb, grec, iseg, irec, nptr, nent;
  size_t bytes() const;
};
template <typename T>
std::shared_ptr<const NfftGeom<T>> nfft_geometry(int N, int n, int W, const double *x, int M);

// Square N x N NFFT with plane batch P (all planes share the nodes).
template <typename T>
struct NfftOp {
  int N = 0, n = 0, W = 0, M = 0, P = 0;
  std::shared_ptr<const NfftGeom<T>> g;  // shared node tables
  bool patched = false;                  // patch-form gather in use (P % 64 == 0)
  bool blocked = false;                  // block-patch interpolation in use (P % 64 == 0)
  DevBuf part;                           // the interpolation's partial sums [entry][P]
  DevBuf gpart;                          // the gather's segment sums [slot][patch cells][P]
  DevBuf X, grid;                        // [N][n][P], [n][n][P]
  size_t bytes() const;
  void init(int N, int n, int W, const double *x, int M, int P);
  // f = A fhat; with sel only at the msel nodes sel[] (other rows untouched)
  void trafo(const cx<T> *fhat, cx<T> *f, hipStream_t st, const int *sel = nullptr, int msel = 0);
  // the R sums of A fhat against meas over the msel measured nodes sel[]:
  // part1/part2[nblocks][P] block sums of |A fhat - m|^2 and |A fhat + m|^2
  void trafo_rsums(const cx<T> *fhat, const cx<T> *meas, const int *sel, int msel, double *part1, double *part2,
                   int nblocks, hipStream_t st);
  void grid_fft(const cx<T> *fhat, hipStream_t st);  // deconvolve, pad, 2D FFT into grid
  // allocate what trafo needs beyond init (the patch form's partial sums; a
  // first trafo does it too, which a graph capture must not contain)
  void prepare_trafo();
  void adjoint(const cx<T> *f, cx<T> *fhat, hipStream_t st);
};

// Spectrum of a Toeplitz kernel embedded in the 2N x 2N circulant: khat
// cx<T>[L][L] stored [column][row], scaled by 1/L^2, and its real part khat_re
// (the kernel is Hermitian, so the spectrum is real).  toeplitz_spectrum()
// builds it once per device, N, nodes and weights and keeps the most recent.
template <typename T>
struct ToepSpectrum {
  DevBuf khat, khat_re;
};
template <typename T>
std::shared_ptr<const ToepSpectrum<T>> toeplitz_spectrum(int N, const double *x, const double *w, int M,
                                                         hipStream_t st);

// Normal operator A^H W A as a 2N x 2N circulant embedding of its Toeplitz
// kernel T(d) = sum_m w_m exp(2 pi i d.x_m).
template <typename T>
struct ToeplitzOp {
  int N = 0, L = 0, P = 0;
  LineFft<T> fft;
  // the kernel spectrum (shared by every job with the same nodes and weights)
  std::shared_ptr<const ToepSpectrum<T>> spec;
  DevBuf X;     // [N][L][P]; for the fused passes [P/pc][N][L][pc]
  int pc = 0;   // planes per chunk of X in the fused passes
  bool pairs = true;  // lane pairs in the 48-point stage (NUDIM_TOEP_PAIRS=0: off)
  // second kernel (init_second) and intermediate for apply_start
  std::shared_ptr<const ToepSpectrum<T>> spec2;
  DevBuf X2;
  size_t bytes() const;
  void init(int N, const double *x, const double *w, int M, int P, hipStream_t st);
  // the kernel of the weights w2[M] (host) as a second operator T2
  void init_second(const double *x, const double *w2, int M, hipStream_t st);
  // z = add + T2 a - T b on the fused passes (fused() and init_second done);
  // zrows (may be null): [N][P] row sums of |z|^2; qpart (may be null, needs
  // init_rform): [L][P] column sums of a^H T3 a (the third kernel's quadratic
  // form of a, by Parseval over the column pass's spectrum of a)
  void apply_start(const cx<T> *a, const cx<T> *b, const cx<T> *add, cx<T> *z, double *zrows, hipStream_t st,
                   double *qpart = nullptr);
  // the kernel of the weights w3[M] (host) as a third operator T3
  std::shared_ptr<const ToepSpectrum<T>> spec3;
  void init_rform(const double *x, const double *w3, int M, hipStream_t st);
  void apply(const cx<T> *p, cx<T> *q, hipStream_t st);
  // fused CGNR step on the fixed-length passes (when fused()): p from z and
  // the old p (see ToepCg), q = T p, part[N][P] = row sums of Re <p, q>
  bool fused() const { return g_use_fft2 && toep_supported(L); }
  void apply_cg(const ToepCg<T> &cg, cx<T> *q, hipStream_t st);
};

// niter sweeps of discrete TV over rows ylo..yhi (clipped by the caller) on
// a [ny][nx][nz] volume held in u0; u1, g0, g1 are volume-sized scratch, and
// g0, g1 must hold zeros outside rows ylo..yhi (clear them once when they are
// allocated).  part_g and part_a hold nb >= tv_blocks(ny, nx, nz) partial
// sums.  The result is left in u0 (returned); u1's rows are scratch.  With
// K > 1 the buffers hold K volumes one after another (ny nx nz elements
// apart), part_g / part_a K x nb sums and sums 2 K: every launch covers all
// K volumes, each with its own norms, exactly as K separate calls would.
int tv_blocks(int ny, int nx, int nz);
template <typename T>
T *tv_sweeps(T *u0, T *u1, T *g0, T *g1, int ny, int nx, int nz, int ylo, int yhi, float lam, int niter,
             double *part_g, double *part_a, double *sums, int nb, hipStream_t st, int K = 1);

// aniso_diffusion_2d (denoise.c:320-378) on a [ny][nx][nz] volume: niter
// in-place Gauss-Seidel sweeps of every z plane, each plane in LDS when it
// fits (k_diffusion_plane), else one launch per hyperplane (k_diffusion_tau);
// both give the same result bit for bit; K volumes of a batch one after another
template <typename T>
void diffusion_sweeps(T *u, int ny, int nx, int nz, int niter, double lam, double ksq, hipStream_t st, int K = 1);

// per-plane reduction helpers
struct Reducer {
  DevBuf part, part2;  // [kRedBlocks][P] doubles
  int P = 0;
  void init(int P);
  template <typename T>
  void dot(const cx<T> *a, const double *w, long long rows, double *out, hipStream_t st);
  template <typename T>
  void redot(const cx<T> *a, const cx<T> *b, long long rows, double *out, hipStream_t st);
  template <typename T>
  void rfactor(const cx<T> *f, const cx<T> *meas, const int *kind, long long rows, double *d1, double *d2,
               hipStream_t st);
};

// CGNR state for P planes
template <typename T>
struct Cgnr {
  int N = 0, M = 0, P = 0, form = 1;
  NfftOp<T> *op = nullptr;
  ToeplitzOp<T> *toep = nullptr;
  const double *w = nullptr;  // device weights [M] or null
  DevBuf z, pv, q, r, v, t;   // z, p, q: [N^2][P]; r, v, t: [M][P]
  DevBuf dot_z, coef, sc;     // [P] each
  DevBuf zz2, alpha, rowpart; // fused Toeplitz steps: [P], [P], [N][P]
  DevBuf colpart;             // column sums of <p, T p>: [2N][P]
  // kept directions: p_k of every step [kept][N^2][P] and alpha_k [kept][P]
  DevBuf pk, alk;
  int kept = 0;
  Reducer red;
  size_t bytes() const;
  void init(int N, int M, int P, int form, NfftOp<T> *op, ToeplitzOp<T> *toep, const double *w);
  // keep the p_k and alpha_k of up to K steps apart, so that a deferred solve
  // leaves f untouched and the caller applies f + sum_k alpha_k p_k once
  void keep_directions(int K);
  // solver_before_loop_complex + nsteps loop_one_step (fhat in/out; f0_zero skips A*0)
  void solve(const cx<T> *y, cx<T> *fhat, int nsteps, bool f0_zero, hipStream_t st);
  // the same on the fused Toeplitz passes with the starting residual given
  // in z (z = A^H W (y - A fhat)).  With defer_last the last step's
  // fhat += alpha p is left to the caller (p in pv, alpha per plane in alpha);
  // with kept directions (nsteps <= kept) every step's update is: fhat is not
  // touched, p_k is in pk + k N^2 P and alpha_k in alk + k P.
  // With z_rows, rowpart holds the [N][P] row sums of |z|^2 (apply_start).
  void solve_from_residual(cx<T> *fhat, int nsteps, hipStream_t st, bool defer_last = false, bool z_rows = false);
};

}  // namespace nudim
