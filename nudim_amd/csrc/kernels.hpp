// kernels.hpp -- NFFT window kernels, CGNR vector kernels, per-plane
// reductions, TV / diffusion stencils and the density constraint.
//
// Layouts (P = planes of the job, innermost everywhere on the Fourier side):
//   Fourier planes     f[y][x][p]      complex   (N x N x P)
//   oversampled grid   g[r][c][p]      complex   (n x n x P)
//   polar samples      s[m][p]         complex   (M x P)
//   real volume        u[y][x][z]      real      (N x N x nz)
#pragma once

#include "common.hpp"

namespace nudim {

constexpr int kThreads = 256;
constexpr int kRedBlocks = 240;  // partial-sum rows of the per-plane reductions

// ---------------------------------------------------------------------------
// NFFT window interpolation (forward transform) and gather adjoint
// ---------------------------------------------------------------------------

// out[m][p] = sum_{t0,t1} psi0[m][t0] psi1[m][t1] g[(c0+t0) mod n][(c1+t1) mod n][p]
// One group of PC = 2^PCs lanes per node (lanes over planes); the window
// weights are uniform over the group and the grid reads of one tap are PC
// contiguous complex values.  With sel the launch covers the M nodes sel[0..M)
// (the other rows of out are left as they are).
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_interp(const cx<T> *__restrict__ g, cx<T> *__restrict__ out, const int *__restrict__ c0,
         const int *__restrict__ c1, const double *__restrict__ psi0, const double *__restrict__ psi1,
         int M, int n, int W, int P, int PCs, const int *__restrict__ sel) {
  const unsigned lb = xcd_block(blockIdx.x, gridDim.x);
  const int PC = 1 << PCs;
  const int i = lb * (kThreads >> PCs) + (threadIdx.x >> PCs);
  const int p = blockIdx.y * PC + (threadIdx.x & (PC - 1));
  if (i >= M) return;
  const int node = sel ? sel[i] : i;
  const int b0 = c0[node], b1 = c1[node];
  const double *a0 = psi0 + (size_t)node * W, *a1 = psi1 + (size_t)node * W;
  const size_t rs = (size_t)n * P;
  cx<T> acc = mk((T)0, (T)0);
  for (int t0 = 0; t0 < W; t0++) {
    int r = b0 + t0;
    r += r < 0 ? n : 0;
    r -= r >= n ? n : 0;
    const cx<T> *row = g + r * rs + p;
    cx<T> racc = mk((T)0, (T)0);
    for (int t1 = 0; t1 < W; t1++) {
      int c = b1 + t1;
      c += c < 0 ? n : 0;
      c -= c >= n ? n : 0;
      racc = racc + (T)a1[t1] * row[(size_t)c * P];
    }
    acc = acc + (T)a0[t0] * racc;
  }
  out[(size_t)node * P + p] = acc;
}

// one node's interpolated value at plane p with the window width W fixed at
// compile time: the node's 2W weights and W column offsets in registers, the
// W grid reads of a row issued together; same summation order as k_interp
template <typename T, int W>
__device__ __forceinline__ cx<T> interp_node(const cx<T> *__restrict__ g, const int *__restrict__ c0,
                                             const int *__restrict__ c1, const double *__restrict__ psi0,
                                             const double *__restrict__ psi1, int node, int n, int P, int p) {
  const int b0 = c0[node], b1 = c1[node];
  T a0[W], a1[W];
  long long col[W];
#pragma unroll
  for (int t = 0; t < W; t++) {
    a0[t] = (T)psi0[(size_t)node * W + t];
    a1[t] = (T)psi1[(size_t)node * W + t];
    int c = b1 + t;
    c += c < 0 ? n : 0;
    c -= c >= n ? n : 0;
    col[t] = (long long)c * P;
  }
  const size_t rs = (size_t)n * P;
  cx<T> acc = mk((T)0, (T)0);
#pragma unroll
  for (int t0 = 0; t0 < W; t0++) {
    int r = b0 + t0;
    r += r < 0 ? n : 0;
    r -= r >= n ? n : 0;
    const cx<T> *row = g + r * rs + p;
    cx<T> v[W];
#pragma unroll
    for (int t1 = 0; t1 < W; t1++) v[t1] = row[col[t1]];
    cx<T> racc = mk((T)0, (T)0);
#pragma unroll
    for (int t1 = 0; t1 < W; t1++) racc = racc + a1[t1] * v[t1];
    acc = acc + a0[t0] * racc;
  }
  return acc;
}

// k_interp with W fixed at compile time (interp_node)
template <typename T, int W>
__global__ void __launch_bounds__(kThreads)
k_interp_w(const cx<T> *__restrict__ g, cx<T> *__restrict__ out, const int *__restrict__ c0,
           const int *__restrict__ c1, const double *__restrict__ psi0, const double *__restrict__ psi1, int M, int n,
           int P, int PCs, const int *__restrict__ sel) {
  const unsigned lb = xcd_block(blockIdx.x, gridDim.x);
  const int PC = 1 << PCs;
  const int i = lb * (kThreads >> PCs) + (threadIdx.x >> PCs);
  const int p = blockIdx.y * PC + (threadIdx.x & (PC - 1));
  if (i >= M) return;
  const int node = sel ? sel[i] : i;
  out[(size_t)node * P + p] = interp_node<T, W>(g, c0, c1, psi0, psi1, node, n, P, p);
}

// The R sums of reconstruction.c:574-582 straight from the interpolation:
// over the M nodes sel[] (all measured), per plane, sum |v - m|^2 and
// |v + m|^2 with v the interpolated value and m the measured one; no v is
// stored.  Blocks stride over the nodes; part1/part2[blockIdx.x][p] get the
// block's FP64 sums (fixed order: node stride, then the node groups).
template <typename T, int W>
__global__ void __launch_bounds__(kThreads)
k_interp_rsums(const cx<T> *__restrict__ g, const cx<T> *__restrict__ meas, const int *__restrict__ c0,
               const int *__restrict__ c1, const double *__restrict__ psi0, const double *__restrict__ psi1, int M,
               int n, int P, int PCs, const int *__restrict__ sel, double *__restrict__ part1,
               double *__restrict__ part2) {
  __shared__ double s1[kThreads], s2[kThreads];
  const int PC = 1 << PCs, NG = kThreads >> PCs;
  const int grp = threadIdx.x >> PCs, lane = threadIdx.x & (PC - 1);
  const int p = blockIdx.y * PC + lane;
  double a1 = 0, a2 = 0;
  for (int i = blockIdx.x * NG + grp; i < M; i += gridDim.x * NG) {
    const int node = sel[i];
    const cx<T> v = interp_node<T, W>(g, c0, c1, psi0, psi1, node, n, P, p);
    const cx<T> m = meas[(size_t)node * P + p];
    const double dx = (double)v.x - (double)m.x, dy = (double)v.y - (double)m.y;
    const double sx = (double)v.x + (double)m.x, sy = (double)v.y + (double)m.y;
    a1 += dx * dx + dy * dy;
    a2 += sx * sx + sy * sy;
  }
  s1[threadIdx.x] = a1;
  s2[threadIdx.x] = a2;
  __syncthreads();
  if (grp == 0) {
    double t1 = 0, t2 = 0;
    for (int k = 0; k < NG; k++) {
      t1 += s1[k * PC + lane];
      t2 += s2[k * PC + lane];
    }
    part1[(size_t)blockIdx.x * P + p] = t1;
    part2[(size_t)blockIdx.x * P + p] = t2;
  }
}

// k_gather with entries decoded eight at a time so that the weight and
// sample reads of eight entries are in flight together.  Same summation
// order as k_gather.
template <typename T, int W>
__global__ void __launch_bounds__(kThreads)
k_gather_w(const cx<T> *__restrict__ s, cx<T> *__restrict__ g, const int *__restrict__ cell_ptr,
           const uint32_t *__restrict__ ent, const double *__restrict__ psi0, const double *__restrict__ psi1,
           int ncell, int P, int PCs) {
  const unsigned lb = xcd_block(blockIdx.x, gridDim.x);
  const int PC = 1 << PCs;
  const int cell = lb * (kThreads >> PCs) + (threadIdx.x >> PCs);
  const int p = blockIdx.y * PC + (threadIdx.x & (PC - 1));
  if (cell >= ncell) return;
  const int e0 = cell_ptr[cell], e1 = cell_ptr[cell + 1];
  cx<T> acc = mk((T)0, (T)0);
  constexpr int B = 8;
  // batches of B entries; entries past e1 re-read the last one (clamped) and
  // are not accumulated, so every read of a batch is unconditional
  for (int e = e0; e < e1; e += B) {
    uint32_t u[B];
#pragma unroll
    for (int j = 0; j < B; j++) u[j] = ent[min(e + j, e1 - 1)];
    double w0[B], w1[B];
    cx<T> sv[B];
#pragma unroll
    for (int j = 0; j < B; j++) {
      int node = (int)(u[j] & 0xFFFFFu), t0 = (int)((u[j] >> 20) & 63u), t1 = (int)(u[j] >> 26);
      w0[j] = psi0[(size_t)node * W + t0];
      w1[j] = psi1[(size_t)node * W + t1];
      sv[j] = s[(size_t)node * P + p];
    }
#pragma unroll
    for (int j = 0; j < B; j++)
      if (e + j < e1) acc = acc + (T)(w0[j] * w1[j]) * sv[j];
  }
  g[(size_t)cell * P + p] = acc;
}

// g[cell][p] = sum over the cell's taps (node, t0, t1) of
//              psi0[node][t0] psi1[node][t1] s[node][p]
// in node order: a deterministic gather, no atomics.  Writes every cell.
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_gather(const cx<T> *__restrict__ s, cx<T> *__restrict__ g, const int *__restrict__ cell_ptr,
         const uint32_t *__restrict__ ent, const double *__restrict__ psi0,
         const double *__restrict__ psi1, int ncell, int W, int P, int PCs) {
  const unsigned lb = xcd_block(blockIdx.x, gridDim.x);
  const int PC = 1 << PCs;
  const int cell = lb * (kThreads >> PCs) + (threadIdx.x >> PCs);
  const int p = blockIdx.y * PC + (threadIdx.x & (PC - 1));
  if (cell >= ncell) return;
  const int e0 = cell_ptr[cell], e1 = cell_ptr[cell + 1];
  cx<T> acc = mk((T)0, (T)0);
  for (int e = e0; e < e1; e++) {
    uint32_t u = ent[e];
    int node = (int)(u & 0xFFFFFu), t0 = (int)((u >> 20) & 63u), t1 = (int)(u >> 26);
    T w = (T)(psi0[(size_t)node * W + t0] * psi1[(size_t)node * W + t1]);
    acc = acc + w * s[(size_t)node * P + p];
  }
  g[(size_t)cell * P + p] = acc;
}

// ---------------------------------------------------------------------------
// per-plane reductions: rows x P arrays, FP64 partial sums [kRedBlocks][P]
// ---------------------------------------------------------------------------

// block (bx, by): plane chunk by (PC planes), rows bx, bx + gridDim.x, ...
// handled by lane groups; partial[bx][p] (and partial2) written.
template <typename T, int MODE>
__global__ void __launch_bounds__(kThreads)
k_reduce(const cx<T> *__restrict__ a, const cx<T> *__restrict__ b, const double *__restrict__ w,
         const int *__restrict__ kind, long long rows, int P, int PCs, double *__restrict__ part,
         double *__restrict__ part2) {
  // MODE 0: sum_k w_k |a|^2 (w may be null)
  // MODE 1: R-factor sums over kind==1 rows: |a-b|^2 and |a+b|^2
  // MODE 2: Re <a, b> = sum Re(conj(a) b)
  __shared__ double sh[kThreads], sh2[kThreads];
  const int PC = 1 << PCs, lanes = kThreads >> PCs;
  const int pl = threadIdx.x & (PC - 1), lr = threadIdx.x >> PCs;
  const int p = blockIdx.y * PC + pl;
  double s = 0, s2 = 0;
  for (long long k = (long long)blockIdx.x * lanes + lr; k < rows; k += (long long)gridDim.x * lanes) {
    cx<T> v = a[k * P + p];
    if constexpr (MODE == 0) {
      double q = (double)v.x * v.x + (double)v.y * v.y;
      s += w ? w[k] * q : q;
    } else if constexpr (MODE == 1) {
      if (kind[k] == 1) {
        cx<T> m = b[k * P + p];
        double dr = (double)v.x - (double)m.x, di = (double)v.y - (double)m.y;
        double sr = (double)v.x + (double)m.x, si = (double)v.y + (double)m.y;
        s += dr * dr + di * di;
        s2 += sr * sr + si * si;
      }
    } else {
      cx<T> m = b[k * P + p];
      s += (double)v.x * m.x + (double)v.y * m.y;
    }
  }
  sh[threadIdx.x] = s;
  if constexpr (MODE == 1) sh2[threadIdx.x] = s2;
  __syncthreads();
  for (int h = lanes >> 1; h > 0; h >>= 1) {
    if (lr < h) {
      sh[threadIdx.x] += sh[threadIdx.x + h * PC];
      if constexpr (MODE == 1) sh2[threadIdx.x] += sh2[threadIdx.x + h * PC];
    }
    __syncthreads();
  }
  if (lr == 0) {
    part[(size_t)blockIdx.x * P + p] = sh[pl];
    if constexpr (MODE == 1) part2[(size_t)blockIdx.x * P + p] = sh2[pl];
  }
}

// sums the nb partial rows per plane in order: out[p] = sum_b part[b][p]
__global__ void k_sum_partials(const double *__restrict__ part, int nb, int P, double *__restrict__ out);

// CGNR scalar steps (NFFT3 solver_loop_one_step_complex, CGNR branch):
//   mode 0: dot_z[p] = s[p]                                 (before_loop)
//   mode 1: alpha[p] = dot_z[p] / s[p]                      (s = dot_v)
//   mode 2: beta[p] = s[p] / dot_z[p]; dot_z[p] = s[p]      (s = new dot_z)
__global__ void k_cg_scalars(double *__restrict__ dot_z, double *__restrict__ coef,
                             const double *__restrict__ s, int P, int mode);

// f += alpha_p * p_  (over rows x P)
template <typename T>
__global__ void k_axpy_planes(cx<T> *__restrict__ f, const cx<T> *__restrict__ x, const double *__restrict__ coef,
                              double sgn, long long total, int P) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  int p = (int)(i % P);
  T a = (T)(sgn * coef[p]);
  cx<T> v = x[i];
  f[i] = f[i] + a * v;
}

// Toeplitz CGNR step, fused: alpha_p = zz[p] / pq[p];  f += alpha p;
// z -= alpha q;  FP64 partial sums of |z|^2 per (row block, plane).
// Block (bx, by): plane chunk by of PC = 2^PCs planes, rows bx, bx+gridDim.x..
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_cg_update(cx<T> *__restrict__ f, const cx<T> *__restrict__ pv, cx<T> *__restrict__ z, const cx<T> *__restrict__ q,
            const double *__restrict__ zz, const double *__restrict__ pq, long long rows, int P, int PCs,
            double *__restrict__ part) {
  __shared__ double sh[kThreads];
  const int PC = 1 << PCs, lanes = kThreads >> PCs;
  const int pl = threadIdx.x & (PC - 1), lr = threadIdx.x >> PCs;
  const int p = blockIdx.y * PC + pl;
  const double alpha = zz[p] / pq[p];
  const T a = (T)alpha, na = (T)(-alpha);
  double s = 0;
  for (long long k = (long long)blockIdx.x * lanes + lr; k < rows; k += (long long)gridDim.x * lanes) {
    const long long i = k * P + p;
    cx<T> pp = pv[i], qq = q[i], ff = f[i], zv = z[i];
    f[i] = ff + a * pp;
    cx<T> zn = zv + na * qq;
    z[i] = zn;
    s += (double)zn.x * zn.x + (double)zn.y * zn.y;
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int h = lanes >> 1; h > 0; h >>= 1) {
    if (lr < h) sh[threadIdx.x] += sh[threadIdx.x + h * PC];
    __syncthreads();
  }
  if (lr == 0) part[(size_t)blockIdx.x * P + p] = sh[pl];
}

// Toeplitz CGNR step without the f update (which the next row pass does):
// alpha_p = zz[p] / pq[p] (kept in alpha_out);  z -= alpha q;  FP64 partial
// sums of |z|^2 per (row block, plane), blocks as in k_cg_update.
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_cg_zupdate(cx<T> *__restrict__ z, const cx<T> *__restrict__ q, const double *__restrict__ zz,
             const double *__restrict__ pq, double *__restrict__ alpha_out, long long rows, int P, int PCs,
             double *__restrict__ part) {
  __shared__ double sh[kThreads];
  const int PC = 1 << PCs, lanes = kThreads >> PCs;
  const int pl = threadIdx.x & (PC - 1), lr = threadIdx.x >> PCs;
  const int p = blockIdx.y * PC + pl;
  const double alpha = zz[p] / pq[p];
  const T na = (T)(-alpha);
  double s = 0;
  for (long long k = (long long)blockIdx.x * lanes + lr; k < rows; k += (long long)gridDim.x * lanes) {
    const long long i = k * P + p;
    cx<T> zn = z[i] + na * q[i];
    z[i] = zn;
    s += (double)zn.x * zn.x + (double)zn.y * zn.y;
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int h = lanes >> 1; h > 0; h >>= 1) {
    if (lr < h) sh[threadIdx.x] += sh[threadIdx.x + h * PC];
    __syncthreads();
  }
  if (lr == 0) {
    part[(size_t)blockIdx.x * P + p] = sh[pl];
    if (blockIdx.x == 0) alpha_out[p] = alpha;
  }
}

// p = beta p + z with beta_p = zz_new[p] / zz_old[p]
template <typename T>
__global__ void k_xpby_ratio(cx<T> *__restrict__ pv, const cx<T> *__restrict__ z, const double *__restrict__ zz_new,
                             const double *__restrict__ zz_old, long long total, int P) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  int p = (int)(i % P);
  T b = (T)(zz_new[p] / zz_old[p]);
  pv[i] = b * pv[i] + z[i];
}

// p = beta_p * p + z
template <typename T>
__global__ void k_xpby_planes(cx<T> *__restrict__ pv, const cx<T> *__restrict__ z, const double *__restrict__ beta,
                              long long total, int P) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  int p = (int)(i % P);
  T b = (T)beta[p];
  pv[i] = b * pv[i] + z[i];
}

// f += alpha_p p ; p = z + beta_p p   (the NFFT-form CG step's two updates of
// p's old value in one pass; the same arithmetic as k_axpy_planes then
// k_xpby_planes)
template <typename T>
__global__ void k_fxp_planes(cx<T> *__restrict__ f, cx<T> *__restrict__ pv, const cx<T> *__restrict__ z,
                             const double *__restrict__ alpha, const double *__restrict__ beta, long long total,
                             int P) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  int p = (int)(i % P);
  const T a = (T)alpha[p], b = (T)beta[p];
  const cx<T> po = pv[i];
  f[i] = f[i] + a * po;
  pv[i] = b * po + z[i];
}

// r -= alpha_p v ; t = w .* r   (rows = M)
template <typename T>
__global__ void k_resid_update(cx<T> *__restrict__ r, const cx<T> *__restrict__ v, const double *__restrict__ alpha,
                               const double *__restrict__ w, cx<T> *__restrict__ t, long long total, int P) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  int p = (int)(i % P);
  long long k = i / P;
  T a = (T)(-alpha[p]);
  cx<T> rv = r[i] + a * v[i];
  r[i] = rv;
  t[i] = w ? (T)w[k] * rv : rv;
}

// r = y - v ; t = w .* r   (before_loop)
template <typename T>
__global__ void k_resid_init(const cx<T> *__restrict__ y, const cx<T> *__restrict__ v, const double *__restrict__ w,
                             cx<T> *__restrict__ r, cx<T> *__restrict__ t, long long total, int P) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  long long k = i / P;
  cx<T> rv = v ? y[i] - v[i] : y[i];
  if (r) r[i] = rv;
  t[i] = w ? (T)w[k] * rv : rv;
}

// t = known ? w m : 0, the weighted known-node data of the projected CG start
template <typename T>
__global__ void k_known_weighted(const cx<T> *__restrict__ meas, const int *__restrict__ kind,
                                 const double *__restrict__ w, cx<T> *__restrict__ t, long long total, int P) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  long long k = i / P;
  t[i] = kind[k] == 1 ? (w ? (T)w[k] * meas[i] : meas[i]) : mk((T)0, (T)0);
}

// Fourier projector, reconstruction.c:574-582: y = known ? measured : f
template <typename T>
__global__ void k_projector(const cx<T> *__restrict__ f, const cx<T> *__restrict__ meas, const int *__restrict__ kind,
                            cx<T> *__restrict__ y, long long total, int P) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  long long k = i / P;
  y[i] = kind[k] == 1 ? meas[i] : f[i];
}

template <typename T>
__global__ void k_copy_cx(cx<T> *__restrict__ d, const cx<T> *__restrict__ s, long long total) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < total) d[i] = s[i];
}

// a -= b
template <typename T>
__global__ void k_sub_cx(cx<T> *__restrict__ a, const cx<T> *__restrict__ b, long long total) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < total) a[i] = a[i] - b[i];
}

template <typename T>
__global__ void k_zero_cx(cx<T> *__restrict__ d, long long total) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < total) d[i] = mk((T)0, (T)0);
}

// ---------------------------------------------------------------------------
// discrete TV (denoise.c:67-249) on u[y][x][z]
// ---------------------------------------------------------------------------

// t / sqrt(ss): exact in FP64 (the reference's arithmetic), t * rsqrt(ss) in FP32
template <typename T>
__device__ __forceinline__ T tv_rs(T t, T ss) {
  if constexpr (sizeof(T) == 4) return t * rsqrtf(ss);
  else return t / sqrt(ss);
}

// Thread walk of the TV kernels over one band row y: flat (x, z) positions
// rem = x nz + z from bxk*kThreads + threadIdx.x in steps of the row's block
// count times kThreads, with x and z carried along (one division per thread)
// and 32-bit byte offsets (the volume holds fewer than 2^32 bytes).
struct TvWalk {
  unsigned rem, x, z, stride, sdx, sdz;
  __device__ TvWalk(unsigned bxk, unsigned nblk, unsigned nz) {
    rem = bxk * kThreads + threadIdx.x;
    x = rem / nz;
    z = rem - x * nz;
    stride = nblk * kThreads;
    sdx = stride / nz;
    sdz = stride - sdx * nz;
  }
  __device__ void next(unsigned nz) {
    rem += stride;
    x += sdx;
    z += sdz;
    if (z >= nz) {
      z -= nz;
      x++;
    }
  }
};

// element at a 32-bit byte offset from p (the global_load form with a scalar
// base and a 32-bit vector offset)
template <typename T>
__device__ __forceinline__ T ld_at(const T *__restrict__ p, unsigned b) {
  return *(const T *)((const char *)p + b);
}
template <typename T>
__device__ __forceinline__ void st_at(T *__restrict__ p, unsigned b, T v) {
  *(T *)((char *)p + b) = v;
}

// grad over rows ylo..yhi (inclusive; clipped), FP64 partial sums of grad^2
// and of u^2 over u < 0 (the pre-update dA, denoise.c:128-133)
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_tv_grad(const T *__restrict__ u, T *__restrict__ grad, int ny, int nx, int nz, int ylo, int yhi,
          double *__restrict__ part_g, double *__restrict__ part_a, long long vs, int pst) {
  // volume blockIdx.z of a batch: volumes vs elements apart, partial sums pst apart
  u += blockIdx.z * vs;
  grad += blockIdx.z * vs;
  part_g += blockIdx.z * pst;
  part_a += blockIdx.z * pst;
  __shared__ double sg[kThreads], sa[kThreads];
  const T eps = (T)1e-8;
  double accg = 0, acca = 0;
  // blocks: (chunk of one row (x, z), band row), renumbered so that each XCD
  // takes a contiguous run of rows (the y neighbours then share its L2)
  const unsigned lb = xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
  const unsigned bxk = lb % gridDim.x;
  const int y = ylo + (int)(lb / gridDim.x);
  constexpr unsigned E = sizeof(T);
  const unsigned unz = (unsigned)nz, rl = (unsigned)nx * unz, sx = unz * E, sy = rl * E;
  const unsigned ybase = (unsigned)y * sy;
  const unsigned dym = y > 0 ? sy : 0u, dyp = y < ny - 1 ? sy : 0u;
  auto at = [&](unsigned b) { return ld_at(u, b); };
  for (TvWalk w(bxk, gridDim.x, unz); w.rem < rl; w.next(unz)) {
    const unsigned o = ybase + w.rem * E;
    const unsigned dzm = w.z > 0 ? E : 0u, dzp = w.z + 1 < unz ? E : 0u;
    const unsigned dxm = w.x > 0 ? sx : 0u, dxp = w.x + 1 < (unsigned)nx ? sx : 0u;
    T c = at(o);
    // reference (i,j,k) = (z,y,x)
    T t1 = c - at(o - dzm);
    T t2 = c - at(o - dym);
    T t3 = c - at(o - dxm);
    T uzp = at(o + dzp);
    T t4 = uzp - c;
    T t5 = uzp - at(o + dzp - dym);
    T t6 = uzp - at(o + dzp - dxm);
    T uyp = at(o + dyp);
    T t7 = uyp - at(o + dyp - dzm);
    T t8 = uyp - c;
    T t9 = uyp - at(o + dyp - dxm);
    T uxp = at(o + dxp);
    T t10 = uxp - at(o + dxp - dzm);
    T t11 = uxp - at(o + dxp - dym);
    T t12 = uxp - c;
    T gv = tv_rs(t1 + t2 + t3, eps + t1 * t1 + t2 * t2 + t3 * t3) - tv_rs(t4, eps + t4 * t4 + t5 * t5 + t6 * t6) -
           tv_rs(t8, eps + t7 * t7 + t8 * t8 + t9 * t9) - tv_rs(t12, eps + t10 * t10 + t11 * t11 + t12 * t12);
    st_at(grad, o, gv);
    accg += (double)gv * (double)gv;
    if (c < (T)0) acca += (double)c * (double)c;
  }
  sg[threadIdx.x] = accg;
  sa[threadIdx.x] = acca;
  __syncthreads();
  for (int h = kThreads / 2; h > 0; h >>= 1) {
    if (threadIdx.x < h) {
      sg[threadIdx.x] += sg[threadIdx.x + h];
      sa[threadIdx.x] += sa[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part_g[lb] = sg[0];
    part_a[lb] = sa[0];
  }
}

// dst = u - lam * sqrt(dA) * (grad / sqrt(sum grad^2))  (denoise.c:125-142)
// over the band rows (dst may be u)
template <typename T>
__global__ void k_tv_update(T *dst, const T *u, const T *__restrict__ grad, int nx, int nz, int ylo, int yhi,
                            const double *__restrict__ sums, double lam, long long vs) {
  // volume blockIdx.y of a batch, its norm pair at sums + 2 blockIdx.y
  dst += blockIdx.y * vs;
  u += blockIdx.y * vs;
  grad += blockIdx.y * vs;
  sums += 2 * blockIdx.y;
  const long long rowlen = (long long)nx * nz;
  const long long total = (long long)(yhi - ylo + 1) * rowlen;
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  long long o = (long long)ylo * rowlen + i;
  const double s1d = sqrt(sums[0]);
  T c = (T)(lam * sqrt(sums[1]));
  T gn;
  if constexpr (sizeof(T) == 4) gn = grad[o] * (T)(1.0 / s1d);
  else gn = grad[o] / s1d;
  dst[o] = u[o] - c * gn;
}

__global__ void k_sum_scalars(const double *__restrict__ part_g, const double *__restrict__ part_a, int nb,
                              double *__restrict__ sums, int pst);

// One fused TV sweep s -> s+1 (denoise.c:83-145 with the next sweep's
// gradient):  reads u_s and grad_s, forms u_{s+1} = u_s - c_s grad_s/|grad_s|
// (c_s = lam sqrt(dA_s)) at the 13 stencil points of each band voxel,
// writes u_{s+1} and grad_{s+1}, and FP64 partial sums of grad_{s+1}^2 and of
// u_{s+1}^2 over u_{s+1} < 0.  Rows outside the band are never updated.
// FP64 keeps the reference's division and square root; FP32 multiplies by
// the reciprocal norm and uses rsqrt.
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_tv_fused(const T *__restrict__ u, const T *__restrict__ g, T *__restrict__ un, T *__restrict__ gn, int ny, int nx,
           int nz, int ylo, int yhi, const double *__restrict__ sums, double lam, double *__restrict__ part_g,
           double *__restrict__ part_a, long long vs, int pst) {
  u += blockIdx.z * vs;
  g += blockIdx.z * vs;
  un += blockIdx.z * vs;
  gn += blockIdx.z * vs;
  sums += 2 * blockIdx.z;
  part_g += blockIdx.z * pst;
  part_a += blockIdx.z * pst;
  __shared__ double sg[kThreads], sa[kThreads];
  const T eps = (T)1e-8;
  const double s1d = sqrt(sums[0]);
  const T s1 = (T)s1d, inv_s1 = (T)(1.0 / s1d);
  const T c = (T)(lam * sqrt(sums[1]));
  // u_{s+1} at flat offset o.  The gradient buffers hold zeros outside the
  // band (they are cleared when allocated and only band rows are written),
  // so the update needs no row test and every stencil read is unconditional.
  auto nxt = [&](unsigned b) -> T {
    T v = ld_at(u, b), gg = ld_at(g, b);
    if constexpr (sizeof(T) == 4) return v - c * (gg * inv_s1);
    else return v - c * (gg / s1);
  };
  double accg = 0, acca = 0;
  // blocks and walk as in k_tv_grad
  const unsigned lb = xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
  const unsigned bxk = lb % gridDim.x;
  const int y = ylo + (int)(lb / gridDim.x);
  constexpr unsigned E = sizeof(T);
  const unsigned unz = (unsigned)nz, rl = (unsigned)nx * unz, sx = unz * E, sy = rl * E;
  const unsigned ybase = (unsigned)y * sy;
  const unsigned dym = y > 0 ? sy : 0u, dyp = y < ny - 1 ? sy : 0u;
  for (TvWalk w(bxk, gridDim.x, unz); w.rem < rl; w.next(unz)) {
    const unsigned o = ybase + w.rem * E;
    const unsigned dzm = w.z > 0 ? E : 0u, dzp = w.z + 1 < unz ? E : 0u;
    const unsigned dxm = w.x > 0 ? sx : 0u, dxp = w.x + 1 < (unsigned)nx ? sx : 0u;
    T cc = nxt(o);
    T t1 = cc - nxt(o - dzm);
    T t2 = cc - nxt(o - dym);
    T t3 = cc - nxt(o - dxm);
    T uzp = nxt(o + dzp);
    T t4 = uzp - cc;
    T t5 = uzp - nxt(o + dzp - dym);
    T t6 = uzp - nxt(o + dzp - dxm);
    T uyp = nxt(o + dyp);
    T t7 = uyp - nxt(o + dyp - dzm);
    T t8 = uyp - cc;
    T t9 = uyp - nxt(o + dyp - dxm);
    T uxp = nxt(o + dxp);
    T t10 = uxp - nxt(o + dxp - dzm);
    T t11 = uxp - nxt(o + dxp - dym);
    T t12 = uxp - cc;
    T gv = tv_rs(t1 + t2 + t3, eps + t1 * t1 + t2 * t2 + t3 * t3) - tv_rs(t4, eps + t4 * t4 + t5 * t5 + t6 * t6) -
           tv_rs(t8, eps + t7 * t7 + t8 * t8 + t9 * t9) - tv_rs(t12, eps + t10 * t10 + t11 * t11 + t12 * t12);
    st_at(un, o, cc);
    st_at(gn, o, gv);
    accg += (double)gv * (double)gv;
    if (cc < (T)0) acca += (double)cc * (double)cc;
  }
  sg[threadIdx.x] = accg;
  sa[threadIdx.x] = acca;
  __syncthreads();
  for (int h = kThreads / 2; h > 0; h >>= 1) {
    if (threadIdx.x < h) {
      sg[threadIdx.x] += sg[threadIdx.x + h];
      sa[threadIdx.x] += sa[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part_g[lb] = sg[0];
    part_a[lb] = sa[0];
  }
}

// k_tv_fused on Q = 4 or 8 consecutive z per lane (nz % Q == 0).  A lane
// reads the seven stencil columns of its run z0..z0+Q-1 with 16-byte loads;
// the values one step outside the run (z0-1 of the centre, y+1 and x+1
// columns, z0+Q of the centre, y-1 and x-1 columns) come from the
// neighbouring lanes, which hold the adjacent runs of the same (y, x) line,
// and are read directly only by a wave's first or last lane when its run
// does not start or end a line (Q = 8 makes a wave cover a whole 512-long
// line).  Each voxel's arithmetic is that of k_tv_fused.
template <typename T, int Q>
struct alignas(16) Run {
  T v[Q];
};

template <typename T, int Q>
__global__ void __launch_bounds__(kThreads)
k_tv_fused_q(const T *__restrict__ u, const T *__restrict__ g, T *__restrict__ un, T *__restrict__ gn, int ny,
             int nx, int nz, int ylo, int yhi, const double *__restrict__ sums, double lam,
             double *__restrict__ part_g, double *__restrict__ part_a, long long vs, int pst) {
  u += blockIdx.z * vs;
  g += blockIdx.z * vs;
  un += blockIdx.z * vs;
  gn += blockIdx.z * vs;
  sums += 2 * blockIdx.z;
  part_g += blockIdx.z * pst;
  part_a += blockIdx.z * pst;
  static_assert(Q == 4 || Q == 8, "runs of 4 or 8");
  __shared__ double sg[kThreads], sa[kThreads];
  const T eps = (T)1e-8;
  const double s1d = sqrt(sums[0]);
  const T s1 = (T)s1d, inv_s1 = (T)(1.0 / s1d);
  const T c = (T)(lam * sqrt(sums[1]));
  auto upd = [&](T v, T gg) -> T {
    if constexpr (sizeof(T) == 4) return v - c * (gg * inv_s1);
    else return v - c * (gg / s1);
  };
  using R = Run<T, Q>;
  auto nxtq = [&](unsigned b, T(&r)[Q]) {
    const R v = ld_at(reinterpret_cast<const R *>(u), b), gg = ld_at(reinterpret_cast<const R *>(g), b);
#pragma unroll
    for (int k = 0; k < Q; k++) r[k] = upd(v.v[k], gg.v[k]);
  };
  auto nxt1 = [&](unsigned b) -> T { return upd(ld_at(u, b), ld_at(g, b)); };
  // the stencil of k_tv_fused at one voxel, neighbour values named as there
  auto voxel = [&](T cv, T zmv, T zpv, T ymv, T ympv, T xmv, T xmpv, T ypv, T ypmv, T ypxmv, T xpv, T xpmv,
                   T ymxpv) -> T {
    T t1 = cv - zmv;
    T t2 = cv - ymv;
    T t3 = cv - xmv;
    T uzp = zpv;
    T t4 = uzp - cv;
    T t5 = uzp - ympv;
    T t6 = uzp - xmpv;
    T uyp = ypv;
    T t7 = uyp - ypmv;
    T t8 = uyp - cv;
    T t9 = uyp - ypxmv;
    T uxp = xpv;
    T t10 = uxp - xpmv;
    T t11 = uxp - ymxpv;
    T t12 = uxp - cv;
    return tv_rs(t1 + t2 + t3, eps + t1 * t1 + t2 * t2 + t3 * t3) - tv_rs(t4, eps + t4 * t4 + t5 * t5 + t6 * t6) -
           tv_rs(t8, eps + t7 * t7 + t8 * t8 + t9 * t9) - tv_rs(t12, eps + t10 * t10 + t11 * t11 + t12 * t12);
  };
  const int lane = threadIdx.x & 63;
  double accg = 0, acca = 0;
  const unsigned lb = xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
  const unsigned bxk = lb % gridDim.x;
  const int y = ylo + (int)(lb / gridDim.x);
  constexpr unsigned E = sizeof(T);
  const unsigned hz = (unsigned)nz / Q, rl = (unsigned)nx * hz, sx = (unsigned)nz * E, sy = (unsigned)nx * sx;
  const unsigned ybase = (unsigned)y * sy;
  const unsigned dym = y > 0 ? sy : 0u, dyp = y < ny - 1 ? sy : 0u;
  for (TvWalk w(bxk, gridDim.x, hz); w.rem < rl; w.next(hz)) {
    const unsigned o = ybase + w.x * sx + w.z * Q * E;
    const unsigned dxm = w.x > 0 ? sx : 0u, dxp = w.x + 1 < (unsigned)nx ? sx : 0u;
    T cc[Q], ym[Q], xm[Q], yp[Q], ypxm[Q], xp[Q], ymxp[Q];
    nxtq(o, cc);
    nxtq(o - dym, ym);
    nxtq(o - dxm, xm);
    nxtq(o + dyp, yp);
    nxtq(o + dyp - dxm, ypxm);
    nxtq(o + dxp, xp);
    nxtq(o + dxp - dym, ymxp);
    // one step below / above the run: the neighbouring lane's run, a direct
    // read at a wave edge inside the line, the run's own end value at the
    // line's ends (the stencil's clamp)
    const bool hm = w.z > 0, hp = w.z + 1 < hz;
    T c_m = __shfl_up(cc[Q - 1], 1), yp_m = __shfl_up(yp[Q - 1], 1), xp_m = __shfl_up(xp[Q - 1], 1);
    T c_p = __shfl_down(cc[0], 1), ym_p = __shfl_down(ym[0], 1), xm_p = __shfl_down(xm[0], 1);
    if (hm && lane == 0) {
      c_m = nxt1(o - E);
      yp_m = nxt1(o + dyp - E);
      xp_m = nxt1(o + dxp - E);
    }
    if (hp && lane == 63) {
      c_p = nxt1(o + Q * E);
      ym_p = nxt1(o - dym + Q * E);
      xm_p = nxt1(o - dxm + Q * E);
    }
    if (!hm) {
      c_m = cc[0];
      yp_m = yp[0];
      xp_m = xp[0];
    }
    if (!hp) {
      c_p = cc[Q - 1];
      ym_p = ym[Q - 1];
      xm_p = xm[Q - 1];
    }
    R gq, cq;
#pragma unroll
    for (int k = 0; k < Q; k++) {
      gq.v[k] = voxel(cc[k], k > 0 ? cc[k - 1] : c_m, k < Q - 1 ? cc[k + 1] : c_p, ym[k],
                      k < Q - 1 ? ym[k + 1] : ym_p, xm[k], k < Q - 1 ? xm[k + 1] : xm_p, yp[k],
                      k > 0 ? yp[k - 1] : yp_m, ypxm[k], xp[k], k > 0 ? xp[k - 1] : xp_m, ymxp[k]);
      cq.v[k] = cc[k];
    }
    st_at(reinterpret_cast<R *>(un), o, cq);
    st_at(reinterpret_cast<R *>(gn), o, gq);
#pragma unroll
    for (int k = 0; k < Q; k++) {
      accg += (double)gq.v[k] * (double)gq.v[k];
      if (cc[k] < (T)0) acca += (double)cc[k] * (double)cc[k];
    }
  }
  sg[threadIdx.x] = accg;
  sa[threadIdx.x] = acca;
  __syncthreads();
  for (int h = kThreads / 2; h > 0; h >>= 1) {
    if (threadIdx.x < h) {
      sg[threadIdx.x] += sg[threadIdx.x + h];
      sa[threadIdx.x] += sa[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part_g[lb] = sg[0];
    part_a[lb] = sa[0];
  }
}

// k_tv_fused_q (runs of 4 z per lane) marching down y: a lane owns one run
// (x, z0..z0+3) over a chunk of R rows of the band and keeps the updated
// columns it will reuse in registers -- (y-1, x) and (y, x) from the steps
// before, (y, x-1) and (y-1, x+1) loaded one step earlier -- so a step loads
// only (y+1, x), (y+1, x-1) and (y, x+1): 3 columns instead of 7, and the
// y+-1 rows are never re-read from L2.  Each voxel's arithmetic is that of
// k_tv_fused.  Grid: (runs / kThreads, chunks); one partial sum per block.
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_tv_march(const T *__restrict__ u, const T *__restrict__ g, T *__restrict__ un, T *__restrict__ gn, int ny, int nx,
           int nz, int ylo, int yhi, int R, const double *__restrict__ sums, double lam, double *__restrict__ part_g,
           double *__restrict__ part_a, long long vs, int pst) {
  u += blockIdx.z * vs;
  g += blockIdx.z * vs;
  un += blockIdx.z * vs;
  gn += blockIdx.z * vs;
  sums += 2 * blockIdx.z;
  part_g += blockIdx.z * pst;
  part_a += blockIdx.z * pst;
  constexpr int Q = 4;
  __shared__ double sg[kThreads], sa[kThreads];
  const T eps = (T)1e-8;
  const double s1d = sqrt(sums[0]);
  const T s1 = (T)s1d, inv_s1 = (T)(1.0 / s1d);
  const T c = (T)(lam * sqrt(sums[1]));
  auto upd = [&](T v, T gg) -> T {
    if constexpr (sizeof(T) == 4) return v - c * (gg * inv_s1);
    else return v - c * (gg / s1);
  };
  using RQ = Run<T, Q>;
  auto nxtq = [&](unsigned b, T(&r)[Q]) {
    const RQ v = ld_at(reinterpret_cast<const RQ *>(u), b), gg = ld_at(reinterpret_cast<const RQ *>(g), b);
#pragma unroll
    for (int k = 0; k < Q; k++) r[k] = upd(v.v[k], gg.v[k]);
  };
  auto nxt1 = [&](unsigned b) -> T { return upd(ld_at(u, b), ld_at(g, b)); };
  auto voxel = [&](T cv, T zmv, T zpv, T ymv, T ympv, T xmv, T xmpv, T ypv, T ypmv, T ypxmv, T xpv, T xpmv,
                   T ymxpv) -> T {
    T t1 = cv - zmv;
    T t2 = cv - ymv;
    T t3 = cv - xmv;
    T uzp = zpv;
    T t4 = uzp - cv;
    T t5 = uzp - ympv;
    T t6 = uzp - xmpv;
    T uyp = ypv;
    T t7 = uyp - ypmv;
    T t8 = uyp - cv;
    T t9 = uyp - ypxmv;
    T uxp = xpv;
    T t10 = uxp - xpmv;
    T t11 = uxp - ymxpv;
    T t12 = uxp - cv;
    return tv_rs(t1 + t2 + t3, eps + t1 * t1 + t2 * t2 + t3 * t3) - tv_rs(t4, eps + t4 * t4 + t5 * t5 + t6 * t6) -
           tv_rs(t8, eps + t7 * t7 + t8 * t8 + t9 * t9) - tv_rs(t12, eps + t10 * t10 + t11 * t11 + t12 * t12);
  };
  const int lane = threadIdx.x & 63;
  double accg = 0, acca = 0;
  const unsigned lb = xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
  const unsigned bx = lb % gridDim.x;
  const int y0 = ylo + (int)(lb / gridDim.x) * R, y1 = min(y0 + R, yhi + 1);
  constexpr unsigned E = sizeof(T);
  const unsigned hz = (unsigned)nz / Q, rl = (unsigned)nx * hz, sx = (unsigned)nz * E, sy = (unsigned)nx * sx;
  const unsigned run = bx * kThreads + threadIdx.x;
  if (run < rl) {
    const unsigned x = run / hz, zr = run - x * hz;
    const unsigned xo = x * sx + zr * Q * E;
    const unsigned dxm = x > 0 ? sx : 0u, dxp = x + 1 < (unsigned)nx ? sx : 0u;
    const bool hm = zr > 0, hp = zr + 1 < hz;
    T cc[Q], ym[Q], xm[Q], ymxp[Q], yp[Q], ypxm[Q], xp[Q];
    {
      const unsigned o = (unsigned)y0 * sy + xo, om = (unsigned)(y0 > 0 ? y0 - 1 : 0) * sy + xo;
      nxtq(o, cc);
      nxtq(om, ym);
      nxtq(o - dxm, xm);
      nxtq(om + dxp, ymxp);
    }
    for (int y = y0; y < y1; y++) {
      const unsigned o = (unsigned)y * sy + xo;
      const unsigned om = (unsigned)(y > 0 ? y - 1 : 0) * sy + xo, op = (unsigned)(y < ny - 1 ? y + 1 : y) * sy + xo;
      nxtq(op, yp);
      nxtq(op - dxm, ypxm);
      nxtq(o + dxp, xp);
      T c_m = __shfl_up(cc[Q - 1], 1), yp_m = __shfl_up(yp[Q - 1], 1), xp_m = __shfl_up(xp[Q - 1], 1);
      T c_p = __shfl_down(cc[0], 1), ym_p = __shfl_down(ym[0], 1), xm_p = __shfl_down(xm[0], 1);
      if (hm && lane == 0) {
        c_m = nxt1(o - E);
        yp_m = nxt1(op - E);
        xp_m = nxt1(o + dxp - E);
      }
      if (hp && lane == 63) {
        c_p = nxt1(o + Q * E);
        ym_p = nxt1(om + Q * E);
        xm_p = nxt1(o - dxm + Q * E);
      }
      if (!hm) {
        c_m = cc[0];
        yp_m = yp[0];
        xp_m = xp[0];
      }
      if (!hp) {
        c_p = cc[Q - 1];
        ym_p = ym[Q - 1];
        xm_p = xm[Q - 1];
      }
      RQ gq, cq;
#pragma unroll
      for (int k = 0; k < Q; k++) {
        gq.v[k] = voxel(cc[k], k > 0 ? cc[k - 1] : c_m, k < Q - 1 ? cc[k + 1] : c_p, ym[k],
                        k < Q - 1 ? ym[k + 1] : ym_p, xm[k], k < Q - 1 ? xm[k + 1] : xm_p, yp[k],
                        k > 0 ? yp[k - 1] : yp_m, ypxm[k], xp[k], k > 0 ? xp[k - 1] : xp_m, ymxp[k]);
        cq.v[k] = cc[k];
      }
      st_at(reinterpret_cast<RQ *>(un), o, cq);
      st_at(reinterpret_cast<RQ *>(gn), o, gq);
#pragma unroll
      for (int k = 0; k < Q; k++) {
        accg += (double)gq.v[k] * (double)gq.v[k];
        if (cc[k] < (T)0) acca += (double)cc[k] * (double)cc[k];
      }
      // the next row's columns
#pragma unroll
      for (int k = 0; k < Q; k++) {
        ym[k] = cc[k];
        cc[k] = yp[k];
        xm[k] = ypxm[k];
        ymxp[k] = xp[k];
      }
    }
  }
  sg[threadIdx.x] = accg;
  sa[threadIdx.x] = acca;
  __syncthreads();
  for (int h = kThreads / 2; h > 0; h >>= 1) {
    if (threadIdx.x < h) {
      sg[threadIdx.x] += sg[threadIdx.x + h];
      sa[threadIdx.x] += sa[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part_g[lb] = sg[0];
    part_a[lb] = sa[0];
  }
}

template <typename T>
__global__ void k_copy_real(T *__restrict__ d, const T *__restrict__ s, long long total, long long vs = 0) {
  d += blockIdx.y * vs;  // volume blockIdx.y of a batch
  s += blockIdx.y * vs;
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < total) d[i] = s[i];
}

// ---------------------------------------------------------------------------
// 2D anisotropic diffusion (denoise.c:320-378) as a hyperplane sweep.
// Update (s, y, x) of sweep s reads the sweep-s values at (y-1,x), (y,x-1)
// and the sweep-(s-1) values at (y+1,x), (y,x+1); all points with
// 2s + y + x == tau are independent, so launch tau = 0, 1, ... in order.
// ---------------------------------------------------------------------------
// Perona-Malik conductance 1 / (1 + g^2 / kappa^2) (denoise.c:357-360):
// the reference's expression in FP64; in FP32 the equal
// kappa^2 / (kappa^2 + g^2), one correctly rounded division instead of two
// (the diffusion call at 128^3: 2.4 -> 1.5 ms)
template <typename T>
__device__ __forceinline__ T pm_conductance(T g, T k2) {
  if constexpr (sizeof(T) == 4) return k2 / (k2 + g * g);
  else return (T)1 / ((T)1 + g * g / k2);
}

// one in-place update of denoise.c:352-369 at offset o of a plane or volume
// (neighbour offsets already clamped)
template <typename T>
__device__ __forceinline__ void pm_update(T *__restrict__ u, long long o, long long dym, long long dyp,
                                          long long dxm, long long dxp, T lam, T k2) {
  T c = u[o];
  T gN = u[o + dym] - c;
  T gS = u[o + dyp] - c;
  T gE = u[o + dxp] - c;
  T gW = u[o + dxm] - c;
  T cN = pm_conductance(gN, k2);
  T cS = pm_conductance(gS, k2);
  T cE = pm_conductance(gE, k2);
  T cW = pm_conductance(gW, k2);
  u[o] = c + lam * (cN * gN + cS * gS + cE * gE + cW * gW);
}

template <typename T>
__global__ void k_diffusion_tau(T *__restrict__ u, int ny, int nx, int nz, int niter, int tau, double lam,
                                double ksq) {
  // threads: (point on the hyperplane) x z ; z fastest
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int z = (int)(i % nz);
  long long j = i / nz;
  // enumerate s in [smin, smax], y in [ymin, ymax] with x = tau - 2s - y
  // j indexes (s, y) pairs with s = smin + j / ny, y = j % ny (invalid x skipped)
  int smin = tau - (ny - 1) - (nx - 1);
  smin = smin > 0 ? (smin + 1) / 2 : 0;
  int s = smin + (int)(j / ny);
  int y = (int)(j % ny);
  if (s >= niter || 2 * s > tau) return;
  int x = tau - 2 * s - y;
  if (x < 0 || x >= nx) return;
  const long long sy = (long long)nx * nz, sx = nz;
  const long long o = (long long)y * sy + (long long)x * sx + z;
  const long long dym = y > 0 ? -sy : 0, dyp = y < ny - 1 ? sy : 0;
  const long long dxm = x > 0 ? -sx : 0, dxp = x < nx - 1 ? sx : 0;
  pm_update<T>(u, o, dym, dyp, dxm, dxp, (T)lam, (T)ksq);
}

// The whole aniso_diffusion_2d call (niter sweeps) for one z plane per
// block, the plane held in LDS: the block loads u[.][.][z], walks the
// hyperplanes tau = 2s + y + x of k_diffusion_tau in order with a block
// barrier between them, and writes the plane back.  The points of one
// hyperplane and their arithmetic are those of k_diffusion_tau, so the
// result is the same bit for bit; HBM sees one read and one write of the
// volume per call instead of one per sweep, and one launch instead of
// 2 (niter - 1) + ny + nx - 1.
template <typename T>
__global__ void __launch_bounds__(1024)
k_diffusion_plane(T *__restrict__ u, int ny, int nx, int nz, int niter, double lam, double ksq) {
  extern __shared__ unsigned char smem_raw[];
  T *pl = reinterpret_cast<T *>(smem_raw);
  // block b: plane z = b mod nz of volume b / nz of a batch
  const int z = blockIdx.x % nz;
  const int nyx = ny * nx;
  u += (long long)(blockIdx.x / nz) * nyx * nz;
  for (int i = threadIdx.x; i < nyx; i += blockDim.x) pl[i] = u[(long long)i * nz + z];
  __syncthreads();
  // thread t takes sweeps s = smin + (t >> 7) + 8 k and rows y = ylo + (t & 127) + 128 k
  // of the valid points (0 <= x = tau - 2s - y < nx) of each hyperplane
  const int sg = threadIdx.x >> 7, l = threadIdx.x & 127;
  const T lm = (T)lam, k2 = (T)ksq;
  const int taumax = 2 * (niter - 1) + (ny - 1) + (nx - 1);
  for (int tau = 0; niter > 0 && tau <= taumax; tau++) {
    int smin = tau - (ny - 1) - (nx - 1);
    smin = smin > 0 ? (smin + 1) / 2 : 0;
    const int smax = min(niter - 1, tau / 2);
    for (int s = smin + sg; s <= smax; s += 8) {
      const int d = tau - 2 * s;  // y + x
      const int ylo = max(0, d - (nx - 1)), yhi = min(ny - 1, d);
      for (int y = ylo + l; y <= yhi; y += 128) {
        const int x = d - y;
        const int o = y * nx + x;
        pm_update<T>(pl, o, y > 0 ? -nx : 0, y < ny - 1 ? nx : 0, x > 0 ? -1 : 0, x < nx - 1 ? 1 : 0, lm, k2);
      }
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < nyx; i += blockDim.x) u[(long long)i * nz + z] = pl[i];
}

// ---------------------------------------------------------------------------
// density constraint (reconstruction.c:672-700) on u[y][x][z]
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256)
k_constraint(T *__restrict__ u, const T *__restrict__ last, int np, int nz, int nx, int is_hio, int config_height,
             int nheight, int ncenter) {
  // grid (rows of np nz values / 256, np, volumes): block y is row yi of
  // volume blockIdx.z, threads run over (x, z) of the row
  const int yi = (int)blockIdx.y, rowlen = np * nz;
  const int r = (int)(blockIdx.x * 256 + threadIdx.x);
  if (r >= rowlen) return;
  const size_t i = ((size_t)blockIdx.z * np + yi) * rowlen + r;
  bool inside;
  if (config_height) {
    inside = yi >= (2 * ncenter - nheight) / 2 && yi <= (2 * ncenter + nheight) / 2;
  } else {
    const int xi = r / nz;
    inside = yi >= (np - nx) / 2 && yi < (np + nx) / 2 && xi >= (np - nx) / 2 && xi < (np + nx) / 2;
  }
  if (!inside && !is_hio) {  // ER outside the support: 0 whatever the value
    u[i] = (T)0;
    return;
  }
  const T v = u[i];
  if (!inside || v < (T)0) u[i] = is_hio ? last[i] - (T)0.9 * v : (T)0;
}

// layout conversions between the reference volume [z][y][x] (double) and
// the device volume [y][x][z] (T)
template <typename T>
__global__ void k_vol_from_ref(const double *__restrict__ ref, T *__restrict__ u, int ny, int nx, int nz) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long total = (long long)ny * nx * nz;
  if (i >= total) return;
  long long yx = i / nz;
  int z = (int)(i % nz);
  u[i] = (T)ref[(long long)z * ny * nx + yx];
}

template <typename T>
__global__ void k_vol_to_ref(const T *__restrict__ u, double *__restrict__ ref, int ny, int nx, int nz) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long total = (long long)ny * nx * nz;
  if (i >= total) return;
  long long z = i / ((long long)ny * nx);
  long long yx = i % ((long long)ny * nx);
  ref[i] = (double)u[yx * nz + z];
}

// planes-major complex (double) [P][K] <-> plane-innermost (T) [K][P]
template <typename T>
__global__ void k_cx_from_planes(const double *__restrict__ src, cx<T> *__restrict__ dst, long long K, int P) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= K * P) return;
  long long k = i / P;
  int p = (int)(i % P);
  const double *s = src + 2 * ((long long)p * K + k);
  dst[i] = mk((T)s[0], (T)s[1]);
}

template <typename T>
__global__ void k_cx_to_planes(const cx<T> *__restrict__ src, double *__restrict__ dst, long long K, int P) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= K * P) return;
  long long p = i / K, k = i % K;
  cx<T> v = src[k * P + p];
  dst[2 * i] = (double)v.x;
  dst[2 * i + 1] = (double)v.y;
}

}  // namespace nudim
