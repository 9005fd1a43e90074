// fft2.hpp -- register-resident four-step line FFTs for fixed lengths.
//
// n = N1 * N2.  With input positions pos = j2 + N2 j1 and output positions
// k1 + N1 k2:
//   X[k1 + N1 k2] = sum_j2 w_N2^{j2 k2} [ w_n^{j2 k1} sum_j1 x[j2 + N2 j1] w_N1^{j1 k1} ]
// Step A: thread (g, j2) loads its N1 inputs straight from HBM, does an
// N1-point DFT in registers and multiplies by w_n^{j2 k1}; one LDS exchange;
// step B: thread (g, k1) does an N2-point DFT in registers and stores.
// The inverse of a transform held as (thread k1, register k2) starts with
// the N2-point DFT over k2 that the thread already holds, so a forward
// transform, a pointwise product and the inverse (the Toeplitz column pass)
// need two LDS exchanges in all.  All index arithmetic inside a line is
// resolved at compile time.
#pragma once

#include "fft.hpp"

namespace nudim {

// ---- compile-time twiddles ---------------------------------------------------

constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double taylor_sin(double x) {
  double x2 = x * x, t = x, s = x;
  for (int i = 1; i < 14; i++) {
    t *= -x2 / ((2.0 * i) * (2.0 * i + 1.0));
    s += t;
  }
  return s;
}
constexpr double taylor_cos(double x) {
  double x2 = x * x, t = 1.0, s = 1.0;
  for (int i = 1; i < 14; i++) {
    t *= -x2 / ((2.0 * i - 1.0) * (2.0 * i));
    s += t;
  }
  return s;
}

struct CS {
  double c, s;
};

// (cos, sin) of 2 pi e / R: octant reduction to |phi| <= pi/4 and an exact
// quarter-turn rotation
constexpr CS cis2pi(long e, long R) {
  e %= R;
  if (e < 0) e += R;
  long o8 = (8 * e) / R;
  double r = (double)(8 * e - o8 * R) / (double)R;  // in [0, 1)
  double c = 0, s = 0;
  long q = 0;
  if (o8 % 2 == 0) {
    double ph = r * (kPi / 4);
    c = taylor_cos(ph);
    s = taylor_sin(ph);
    q = o8 / 2;
  } else {
    double ph = (1.0 - r) * (kPi / 4);  // angle = (o8+1) pi/4 - ph
    c = taylor_cos(ph);
    s = -taylor_sin(ph);
    q = (o8 + 1) / 2;
  }
  switch (q % 4) {
    case 0: return CS{c, s};
    case 1: return CS{-s, c};
    case 2: return CS{-c, -s};
    default: return CS{s, -c};
  }
}

// The R roots of unity exp(S 2 pi i e / R) of one DFT length in the
// transform's precision, evaluated while compiling.  A function-local
// constexpr table is a private constant, so a read at an index known after
// unrolling folds to an immediate; evaluating cis2pi in the kernel body
// instead leaves FP64 series code in it once the unrolled DFT is large.
template <typename T, int S, int R>
struct TwTable {
  T c[R], s[R];
  constexpr TwTable() : c(), s() {
    for (int e = 0; e < R; e++) {
      CS v = cis2pi(e, R);
      c[e] = (T)v.c;
      s[e] = (T)(S * v.s);
    }
  }
};

// w_R^e = exp(S 2 pi i e / R)
template <typename T, int S, int R>
__device__ __forceinline__ cx<T> wconst(int e) {
  constexpr TwTable<T, S, R> tab{};
  const int i = ((e % R) + R) % R;
  return mk(tab.c[i], tab.s[i]);
}

// v * w_R^e with the quarter-turn multiples as exact sign swaps (the same
// values as the product with the table's exact 0 and 1); e is a constant
// once the caller's loops unroll, so the choice is made while compiling
template <typename T, int S, int R>
__device__ __forceinline__ cx<T> mul_wconst(cx<T> v, int e) {
  const int i = ((e % R) + R) % R;
  if (i == 0) return v;
  if (4 * i == R) return S > 0 ? mk(-v.y, v.x) : mk(v.y, -v.x);      // S i
  if (2 * i == R) return mk(-v.x, -v.y);                             // -1
  if (4 * i == 3 * R) return S > 0 ? mk(v.y, -v.x) : mk(-v.y, v.x);  // -S i
  return v * wconst<T, S, R>(e);
}

constexpr int split_factor(int R) {
  return (R % 8 == 0 && R != 8) ? 8 : (R % 4 == 0 && R != 4) ? 4 : (R % 2 == 0 && R != 2) ? 2
       : (R % 3 == 0 && R != 3) ? 3 : (R % 5 == 0 && R != 5) ? 5 : R;
}

// in-register DFT of R points, natural order in and out, sign S
template <typename T, int R, int S>
__device__ __forceinline__ void dft_reg(cx<T> *v) {
  if constexpr (R == 1) {
    return;
  } else if constexpr (R == 2 || R == 3 || R == 4 || R == 5 || R == 8) {
    dft_small<T, R>(v, S);
  } else {
    constexpr int A = split_factor(R);
    constexpr int B = R / A;
    static_assert(A > 1 && B > 1 && A * B == R, "length must factor into 2, 3, 4, 5, 8");
    cx<T> t[R];
#pragma unroll
    for (int j2 = 0; j2 < B; j2++) {
      cx<T> a[A];
#pragma unroll
      for (int j1 = 0; j1 < A; j1++) a[j1] = v[j2 + B * j1];
      dft_reg<T, A, S>(a);
#pragma unroll
      for (int k1 = 0; k1 < A; k1++) t[j2 + B * k1] = mul_wconst<T, S, R>(a[k1], j2 * k1);
    }
#pragma unroll
    for (int k1 = 0; k1 < A; k1++) {
      cx<T> b[B];
#pragma unroll
      for (int j2 = 0; j2 < B; j2++) b[j2] = t[j2 + B * k1];
      dft_reg<T, B, S>(b);
#pragma unroll
      for (int k2 = 0; k2 < B; k2++) v[k1 + A * k2] = b[k2];
    }
  }
}

// ---- pruned DFTs and the LDS exchange ---------------------------------------------

// R-point DFT (sign S) of v whose inputs v[R/2..R) are zero (not read):
// even outputs = DFT_{R/2}(v), odd outputs = DFT_{R/2}(v[j] w_R^{Sj}).
template <typename T, int R, int S>
__device__ __forceinline__ void dft_half_in(cx<T> *v) {
  constexpr int H = R / 2;
  cx<T> e[H], o[H];
#pragma unroll
  for (int j = 0; j < H; j++) {
    e[j] = v[j];
    o[j] = mul_wconst<T, S, R>(v[j], j);
  }
  dft_reg<T, H, S>(e);
  dft_reg<T, H, S>(o);
#pragma unroll
  for (int k = 0; k < H; k++) {
    v[2 * k] = e[k];
    v[2 * k + 1] = o[k];
  }
}

// First R/2 outputs of the R-point DFT (sign S) of v:
// Y[m] = DFT_{R/2}(v[2j])[m] + w_R^{Sm} DFT_{R/2}(v[2j+1])[m].
template <typename T, int R, int S>
__device__ __forceinline__ void dft_half_out(cx<T> *v) {
  constexpr int H = R / 2;
  cx<T> e[H], o[H];
#pragma unroll
  for (int j = 0; j < H; j++) {
    e[j] = v[2 * j];
    o[j] = v[2 * j + 1];
  }
  dft_reg<T, H, S>(e);
  dft_reg<T, H, S>(o);
#pragma unroll
  for (int m = 0; m < H; m++) v[m] = e[m] + mul_wconst<T, S, R>(o[m], m);
}

// Exchange of the four-step transform through LDS.  SPLIT: through a real
// buffer, real parts then imaginary parts -- half the LDS of a complex
// buffer (more blocks per CU) for two more barriers.  On entry the buffer
// must be free; on exit it is still being read.
template <typename T, int NW, int NR, bool SPLIT>
__device__ __forceinline__ void xchg(T *lx, const cx<T> *wv, bool wact, int wbase, int wstride, cx<T> *rv, bool ract,
                                     int rbase, int rstride) {
  if constexpr (!SPLIT) {
    cx<T> *lc = reinterpret_cast<cx<T> *>(lx);
    if (wact) {
#pragma unroll
      for (int k = 0; k < NW; k++) lc[wbase + k * wstride] = wv[k];
    }
    __syncthreads();
    if (ract) {
#pragma unroll
      for (int j = 0; j < NR; j++) rv[j] = lc[rbase + j * rstride];
    }
    return;
  }
  if (wact) {
#pragma unroll
    for (int k = 0; k < NW; k++) lx[wbase + k * wstride] = wv[k].x;
  }
  __syncthreads();
  if (ract) {
#pragma unroll
    for (int j = 0; j < NR; j++) rv[j].x = lx[rbase + j * rstride];
  }
  __syncthreads();
  if (wact) {
#pragma unroll
    for (int k = 0; k < NW; k++) lx[wbase + k * wstride] = wv[k].y;
  }
  __syncthreads();
  if (ract) {
#pragma unroll
    for (int j = 0; j < NR; j++) rv[j].y = lx[rbase + j * rstride];
  }
}

// xchg with the LDS index of each write k and read j given by wi(k), ri(j)
template <typename T, int NW, int NR, bool SPLIT, typename WF, typename RF>
__device__ __forceinline__ void xchg_at(T *lx, const cx<T> *wv, bool wact, WF wi, cx<T> *rv, bool ract, RF ri) {
  if constexpr (!SPLIT) {
    cx<T> *lc = reinterpret_cast<cx<T> *>(lx);
    if (wact) {
#pragma unroll
      for (int k = 0; k < NW; k++) lc[wi(k)] = wv[k];
    }
    __syncthreads();
    if (ract) {
#pragma unroll
      for (int j = 0; j < NR; j++) rv[j] = lc[ri(j)];
    }
    return;
  }
  if (wact) {
#pragma unroll
    for (int k = 0; k < NW; k++) lx[wi(k)] = wv[k].x;
  }
  __syncthreads();
  if (ract) {
#pragma unroll
    for (int j = 0; j < NR; j++) rv[j].x = lx[ri(j)];
  }
  __syncthreads();
  if (wact) {
#pragma unroll
    for (int k = 0; k < NW; k++) lx[wi(k)] = wv[k].y;
  }
  __syncthreads();
  if (ract) {
#pragma unroll
    for (int j = 0; j < NR; j++) rv[j].y = lx[ri(j)];
  }
}

// v * exp(-S 2 pi i q / 4): a quarter-turn rotation (no flops)
template <typename T, int S>
__device__ __forceinline__ cx<T> mul_quarter(cx<T> v, int q) {
  q &= 3;
  if (S > 0) q = (4 - q) & 3;
  switch (q) {
    case 0: return v;
    case 1: return mk(-v.y, v.x);
    case 2: return mk(-v.x, -v.y);
    default: return mk(v.y, -v.x);
  }
}

// ---- the line kernel ------------------------------------------------------------

// One transform of sign S per line.  Threads: N1 * G (>= N2 * G).  LDS: the
// exchange buffer [N2][N1*G + pad] followed by the n twiddles w_n^j.
//
// HIN: the input holds n/2 elements centred on position 0 (cnt = n/2,
// shift = -n/4), i.e. positions j1 in [3N1/4, N1) and [0, N1/4) of step A.
// With b[j'] = a[(j' - N1/4) mod N1], X[k] = exp(-S 2 pi i k/4) DFT(b)[k]
// and b has its upper half zero: a half-size DFT and a quarter turn.
// HOUT: only the n/2 outputs centred on position 0 are stored, i.e. k2 in
// [3N2/4, N2) and [0, N2/4) of step B: Y[m' - N2/4] is output m' < N2/2 of
// the DFT of v[j] exp(-S 2 pi i j/4).
//
// H = 2 (HIN only): two adjacent lanes share line position r = j2 and split
// the half-input N1-point DFT by output parity, as in the Toeplitz passes:
// lane h computes outputs k1 = 2s + h; blocks of max(N1, 2 N2) * G threads.
template <typename T, int N1, int N2, int S, bool HIN, bool HOUT, int H>
__global__ void __launch_bounds__(H == 2 ? 512 : 256)
k_fft2(const cx<T> *in, cx<T> *out, LineIO li, LineIO lo, const cx<T> *__restrict__ tw, int Gs, int nchunk) {
  constexpr int n = N1 * N2, Q1 = N1 / 4, Q2 = N2 / 4;
  constexpr int A = N1 / H;  // step-A values per lane
  static_assert(!HIN || N1 % 4 == 0, "centred input needs N1 % 4 == 0");
  static_assert(!HOUT || N2 % 4 == 0, "centred output needs N2 % 4 == 0");
  static_assert(H == 1 || (H == 2 && HIN), "lane pairs for the half-input form only");
  extern __shared__ unsigned char smem_raw[];
  cx<T> *lds = reinterpret_cast<cx<T> *>(smem_raw);
  const int G = 1 << Gs, G1 = G - 1;
  // exchange row of one j2 (padded by G so that the step-A writes of a
  // wavefront spread evenly over the LDS banks)
  const int rowA = N1 * G + G;
  cx<T> *stw = lds + N2 * rowA;         // n twiddles
  const int lb = (int)xcd_block(blockIdx.x, gridDim.x);
  const int o = lb / nchunk, chunk = lb - o * nchunk;
  const int tid = threadIdx.x;
  for (int i = tid; i < n; i += blockDim.x) stw[i] = tw[i];
  // step A: lane h of line position r = j2, plane g; step B: k1 = kb, plane gb
  const int h = H == 2 ? (tid & 1) : 0;
  const int g = (tid >> (H - 1)) & G1, r = tid >> (Gs + H - 1);
  const int gb = tid & G1, kb = tid >> Gs;
  const bool actA = r < N2, actB = kb < N1;
  auto k1of = [&](int s) { return H == 1 ? s : 2 * s + h; };
  const cx<T> *src = in + (long long)o * li.so + (long long)chunk * G + g;
  const T so_scale = li.scale_o ? (T)li.scale_o[o] : (T)1;

  // ---- step A: load positions j2 + N2 j1, N1-point DFT ----
  cx<T> a[A];
  if (actA) {
    if constexpr (HIN) {
      // every loaded position is in range: all N1/2 loads in flight together
      int kk[N1 / 2];
#pragma unroll
      for (int jp = 0; jp < N1 / 2; jp++)
        kk[jp] = (jp < Q1) ? r + N2 * (jp + 3 * Q1) + n / 4 - n : r + N2 * (jp - Q1) + n / 4;
#pragma unroll
      for (int jp = 0; jp < N1 / 2; jp++) a[jp] = src[(long long)kk[jp] * li.sk];
      if (li.scale_k) {
        T sc[N1 / 2];
#pragma unroll
        for (int jp = 0; jp < N1 / 2; jp++) sc[jp] = (T)li.scale_k[kk[jp]];
#pragma unroll
        for (int jp = 0; jp < N1 / 2; jp++) a[jp] = (sc[jp] * so_scale) * a[jp];
      } else if (li.scale_o) {
#pragma unroll
        for (int jp = 0; jp < N1 / 2; jp++) a[jp] = so_scale * a[jp];
      }
      if constexpr (H == 1) {
        dft_half_in<T, N1, S>(a);
      } else {
        if (h) {
#pragma unroll
          for (int j = 1; j < A; j++) a[j] = mul_wconst<T, S, N1>(a[j], j);
        }
        dft_reg<T, A, S>(a);
      }
#pragma unroll
      for (int s = 0; s < A; s++) a[s] = mul_quarter<T, S>(a[s], k1of(s));
    } else {
      // Loads are unconditional (index clamped into range, out-of-range lanes
      // zeroed afterwards) so that all N1 of them are in flight together.
      int kk[N1];
#pragma unroll
      for (int j1 = 0; j1 < N1; j1++) kk[j1] = wrap_pos(r + N2 * j1 - li.shift, n);
#pragma unroll
      for (int j1 = 0; j1 < N1; j1++) a[j1] = src[(long long)min(kk[j1], li.cnt - 1) * li.sk];
      if (li.scale_k) {
        T sc[N1];
#pragma unroll
        for (int j1 = 0; j1 < N1; j1++) sc[j1] = (T)li.scale_k[min(kk[j1], li.cnt - 1)];
#pragma unroll
        for (int j1 = 0; j1 < N1; j1++) a[j1] = (sc[j1] * so_scale) * a[j1];
      } else if (li.scale_o) {
#pragma unroll
        for (int j1 = 0; j1 < N1; j1++) a[j1] = so_scale * a[j1];
      }
#pragma unroll
      for (int j1 = 0; j1 < N1; j1++)
        if (kk[j1] >= li.cnt) a[j1] = mk((T)0, (T)0);
      dft_reg<T, N1, S>(a);
    }
  }
  __syncthreads();  // twiddle table complete
  if (actA) {
#pragma unroll
    for (int s = 0; s < A; s++) {
      const int k1 = k1of(s);
      cx<T> w = stw[r * k1];
      if (S > 0) w = conjx(w);
      lds[r * rowA + k1 * G + g] = a[s] * w;
    }
  }
  __syncthreads();

  // ---- step B: N2-point DFT over j2 for fixed k1 = kb ----
  if (!actB) return;
  cx<T> b[N2];
#pragma unroll
  for (int j2 = 0; j2 < N2; j2++) b[j2] = lds[j2 * rowA + kb * G + gb];
  cx<T> *dst = out + (long long)o * lo.so + (long long)chunk * G + gb;
  const T oo_scale = lo.scale_o ? (T)lo.scale_o[o] : (T)1;
  constexpr int NO = HOUT ? N2 / 2 : N2;
  int kk[NO];
  if constexpr (HOUT) {
#pragma unroll
    for (int j = 0; j < N2; j++) b[j] = mul_quarter<T, S>(b[j], j);
    dft_half_out<T, N2, S>(b);
#pragma unroll
    for (int mp = 0; mp < NO; mp++)
      kk[mp] = (mp < Q2) ? kb + N1 * (mp + 3 * Q2) + n / 4 - n : kb + N1 * (mp - Q2) + n / 4;
  } else {
    dft_reg<T, N2, S>(b);
#pragma unroll
    for (int k2 = 0; k2 < NO; k2++) kk[k2] = wrap_pos(kb + N1 * k2 - lo.shift, n);
  }
  if (lo.scale_k) {
    T sc[NO];
#pragma unroll
    for (int k2 = 0; k2 < NO; k2++) sc[k2] = (T)lo.scale_k[min(kk[k2], lo.cnt - 1)];
#pragma unroll
    for (int k2 = 0; k2 < NO; k2++) b[k2] = (sc[k2] * oo_scale) * b[k2];
  } else if (lo.scale_o) {
#pragma unroll
    for (int k2 = 0; k2 < NO; k2++) b[k2] = oo_scale * b[k2];
  }
#pragma unroll
  for (int k2 = 0; k2 < NO; k2++)
    if (HOUT || kk[k2] < lo.cnt) dst[(long long)kk[k2] * lo.sk] = b[k2];
}

// ---- lane-pair line kernel (FP64) ----------------------------------------------------
//
// In FP64 a lane holding all N1 = 32 values of a step-A line position (and
// the DFT's temporaries) needs more than 256 registers, and the complex
// exchange of 8 planes 148 KB of LDS: one wavefront per SIMD.  Here two
// adjacent lanes share each step-A line position and each step-B column k1
// and hold half of its values; one radix-2 step across the pair, the values
// swapped with a lane shuffle, splits each DFT into two of half length by
// output parity:
//   X[2k]     = DFT_{R/2}(x[i] + x[i + R/2])[k]                (lane 0)
//   X[2k + 1] = DFT_{R/2}((x[i] - x[i + R/2]) w_R^{S i})[k]    (lane 1)
// with lane h holding inputs i + h R/2 (i < R/2).  The exchange goes through
// a real buffer (real parts, then imaginary parts).

// v[i] = input i + h R/2 on entry; on return v[s] = output 2 s + h
template <typename T, int R, int S>
__device__ __forceinline__ void dft_pair(cx<T> *v, int h) {
  constexpr int Q = R / 2;
#pragma unroll
  for (int i = 0; i < Q; i++) {
    cx<T> o;
    o.x = __shfl_xor(v[i].x, 1);
    o.y = __shfl_xor(v[i].y, 1);
    const cx<T> d = h ? o - v[i] : v[i] + o;
    v[i] = h ? mul_wconst<T, S, R>(d, i) : d;
  }
  dft_reg<T, Q, S>(v);
}

// One transform of sign S per line, full input and output (positions past
// cnt are zero on input and not stored on output).  Threads 2 max(N1, N2) G:
// lane h = tid & 1, plane g, line position r (step A) or k1 (step B) =
// tid >> (log2 G + 1).  LDS: the real exchange buffer [N2][N1 G + G] and the
// n twiddles.
template <typename T, int N1, int N2, int S>
__global__ void __launch_bounds__(256)
k_fft2p(const cx<T> *in, cx<T> *out, LineIO li, LineIO lo, const cx<T> *__restrict__ tw, int Gs, int nchunk) {
  constexpr int n = N1 * N2, A1 = N1 / 2, A2 = N2 / 2;
  static_assert(N1 % 2 == 0 && N2 % 2 == 0, "even factors");
  extern __shared__ unsigned char smem_raw[];
  T *lx = reinterpret_cast<T *>(smem_raw);
  const int G = 1 << Gs, G1 = G - 1;
  const int rowA = N1 * G + G;
  cx<T> *stw = reinterpret_cast<cx<T> *>(lx + N2 * rowA);  // N2 rowA is even: 16-byte aligned
  const int lb = (int)xcd_block(blockIdx.x, gridDim.x);
  const int o = lb / nchunk, chunk = lb - o * nchunk;
  const int tid = threadIdx.x;
  for (int i = tid; i < n; i += blockDim.x) stw[i] = tw[i];
  const int h = tid & 1, g = (tid >> 1) & G1, r = tid >> (Gs + 1);
  const bool actA = r < N2, actB = r < N1;

  // ---- step A: positions r + N2 (h A1 + i), N1-point DFT over the pair ----
  cx<T> a[A1];
  if (actA) {
    const cx<T> *src = in + (long long)o * li.so + (long long)chunk * G + g;
    const T so_scale = li.scale_o ? (T)li.scale_o[o] : (T)1;
    int kk[A1];
#pragma unroll
    for (int i = 0; i < A1; i++) kk[i] = wrap_pos(r + N2 * (h * A1 + i) - li.shift, n);
#pragma unroll
    for (int i = 0; i < A1; i++) {
      cx<T> v = mk((T)0, (T)0);
      if (kk[i] < li.cnt) v = src[(long long)kk[i] * li.sk];
      a[i] = v;
    }
    if (li.scale_k) {
#pragma unroll
      for (int i = 0; i < A1; i++) {
        const T sc = kk[i] < li.cnt ? (T)li.scale_k[kk[i]] : (T)0;
        a[i] = (sc * so_scale) * a[i];
      }
    } else if (li.scale_o) {
#pragma unroll
      for (int i = 0; i < A1; i++) a[i] = so_scale * a[i];
    }
    dft_pair<T, N1, S>(a, h);
  }
  __syncthreads();  // twiddle table complete
  if (actA) {
#pragma unroll
    for (int s = 0; s < A1; s++) {
      cx<T> w = stw[r * (2 * s + h)];
      if (S > 0) w = conjx(w);
      a[s] = a[s] * w;
    }
  }
  // ---- step B: k1 = r, lane h holds j2 = h A2 + i ----
  cx<T> b[A2];
  xchg_at<T, A1, A2, true>(
      lx, a, actA, [&](int s) { return r * rowA + (2 * s + h) * G + g; }, b, actB,
      [&](int i) { return (h * A2 + i) * rowA + r * G + g; });
  if (!actB) return;
  dft_pair<T, N2, S>(b, h);
  cx<T> *dst = out + (long long)o * lo.so + (long long)chunk * G + g;
  const T oo_scale = lo.scale_o ? (T)lo.scale_o[o] : (T)1;
#pragma unroll
  for (int s = 0; s < A2; s++) {
    const int kk = wrap_pos(r + N1 * (2 * s + h) - lo.shift, n);
    if (kk < lo.cnt) {
      cx<T> v = b[s];
      if (lo.scale_k) v = ((T)lo.scale_k[kk] * oo_scale) * v;
      else if (lo.scale_o) v = oo_scale * v;
      dst[(long long)kk * lo.sk] = v;
    }
  }
}

// ---- Toeplitz passes ---------------------------------------------------------------
//
// The circulant embedding of an N x N plane into n x n, n = 2N, holds the
// plane in the first N positions of every line and keeps only the first N
// outputs.  With positions j2 + N2 j1 this makes inputs j1 >= N1/2 zero in
// step A and leaves only outputs m < R/2 wanted in the last DFT, so those
// DFTs run at half size.  The row passes also carry the CGNR vector work.

// Line o of a pass, plane chunk*G + g; element k of a line at
// o*so + k*sk (+ plane).  ROWF: o = row, in = p (N positions) -> out = X
// (n positions).  COL: o = column of X, rows 0..N-1 -> forward, times the
// real mid[o*n + k], inverse, rows 0..N-1 back in place.  ROWI: o = row of X
// (n positions) -> out = q (N positions).  LDS: the exchange buffer
// [N2][N1*G + G] (real with toep_split(KIND), else complex) followed by the
// n twiddles.
//
// H lanes per step-A line.  With H = 2 the two adjacent lanes h = 0, 1 of a
// pair share line position r and each holds N1/2 of its values, which halves
// the registers of the N1-point stage (N1 = 48 at n = 1536):
//   - half-input forward DFT (ROWF, COL): even outputs = DFT_{N1/2}(v), odd
//     outputs = DFT_{N1/2}(v[j] w_N1^{-j}); lane h computes parity h, k1 =
//     2s + h, both lanes loading the N1/2 inputs;
//   - full-input inverse DFT (ROWI): lane h transforms the inputs of parity h
//     (E, O), then X[k] = E[k] + w^k O[k] and X[k + N1/2] = E[k] - w^k O[k]
//     with the lanes swapping N1/4 values: lane h keeps k in [h N1/4,
//     (h+1) N1/4);
//   - half-output last DFT of the column pass: Y[m] = E[m] + w^m O[m] for
//     m < N1/2, E from the even k1, O from the odd, the same swap.
// Step B keeps one thread per (k1, plane): blocks of max(N1, H N2) * G.
constexpr bool toep_split(int kind) { return kind != TK_ROWF; }

// the pair swap of the H = 2 forms: a[0..2Q) holds E (lane 0) or w^k O
// (lane 1) for k < 2Q; on return a[i] = X[k], a[Q + i] = X[k + 2Q] (first
// half-output form: a[i] = Y[k] only) for this lane's k = Q h + i
template <typename T, int Q, bool BOTH>
__device__ __forceinline__ void pair_swap(cx<T> *a, int h) {
  cx<T> c[2 * Q];
#pragma unroll
  for (int i = 0; i < Q; i++) {
    const cx<T> send = h ? a[i] : a[Q + i];
    cx<T> recv;
    recv.x = __shfl_xor(send.x, 1);
    recv.y = __shfl_xor(send.y, 1);
    const cx<T> e = h ? recv : a[i], od = h ? a[Q + i] : recv;
    c[i] = e + od;
    if (BOTH) c[Q + i] = e - od;
  }
#pragma unroll
  for (int i = 0; i < (BOTH ? 2 * Q : Q); i++) a[i] = c[i];
}

// GSC >= 0: log2 G fixed at compile time (the launch's Gs argument is then
// GSC), so that every LDS address of the exchanges is a constant offset
template <typename T, int N1, int N2, int KIND, int H, int GSC>
__global__ void __launch_bounds__((KIND == TK_ROWF || KIND == TK_ROWI) ? 512 : 256)
k_toep(const cx<T> *in, cx<T> *out, long long in_so, long long in_sk, long long out_so, long long out_sk,
       const cx<T> *__restrict__ tw, int Gs_arg, int nchunk, const T *__restrict__ mid, ToepCg<T> cg, int pstride) {
  const int Gs = GSC >= 0 ? GSC : Gs_arg;
  constexpr int n = N1 * N2, H1 = N1 / 2, H2 = N2 / 2;
  constexpr int A = N1 / H;   // step-A values per lane
  constexpr int Q4 = N1 / 4;  // values a pair swaps (H = 2)
  constexpr bool SPLIT = toep_split(KIND);
  static_assert(N1 % 2 == 0 && N2 % 2 == 0 && N1 >= N2, "even factors, N1 >= N2");
  static_assert(H == 1 || (H == 2 && N1 % 4 == 0), "lane pairs need N1 % 4 == 0");
  extern __shared__ unsigned char smem_raw[];
  T *lx = reinterpret_cast<T *>(smem_raw);
  const int G = 1 << Gs, G1 = G - 1;
  const int rowA = N1 * G + G;
  cx<T> *stw = reinterpret_cast<cx<T> *>(lx + (SPLIT ? 1 : 2) * N2 * rowA);
  const int lb = (int)xcd_block(blockIdx.x, gridDim.x);
  const int o = lb / nchunk, chunk = lb - o * nchunk;
  const int tid = threadIdx.x;
  for (int i = tid; i < n; i += blockDim.x) stw[i] = tw[i];
  // step A: lane h of line position r < N2, plane g
  const int h = H == 2 ? (tid & 1) : 0;
  const int g = (tid >> (H - 1)) & G1, r = tid >> (Gs + H - 1);
  const bool actA = r < N2;
  const int plane = chunk * G + g;
  // step B: column k1 = kb < N1, plane gb
  const int gb = tid & G1, kb = tid >> Gs;
  const bool actB = H == 1 ? true : kb < N1;
  const int planeB = chunk * G + gb;
  constexpr int SA = (KIND == TK_ROWI) ? +1 : -1;
  // k1 of step-A value s of this lane
  auto k1of = [&](int s) -> int {
    if constexpr (H == 1) return s;
    else if constexpr (KIND == TK_ROWI) return s < Q4 ? Q4 * h + s : H1 + Q4 * h + (s - Q4);
    else return 2 * s + h;
  };

  // ---- step A ----
  cx<T> a[A];
  if (actA) {
    if constexpr (KIND == TK_ROWI) {
      const cx<T> *src = in + (long long)o * in_so + plane;
      if constexpr (H == 1) {
#pragma unroll
        for (int j1 = 0; j1 < N1; j1++) a[j1] = src[(long long)(r + N2 * j1) * in_sk];
        dft_reg<T, N1, +1>(a);
      } else {
#pragma unroll
        for (int j = 0; j < A; j++) a[j] = src[(long long)(r + N2 * (2 * j + h)) * in_sk];
        dft_reg<T, A, +1>(a);
        if (h) {
#pragma unroll
          for (int k = 1; k < A; k++) a[k] = mul_wconst<T, +1, N1>(a[k], k);
        }
        pair_swap<T, Q4, true>(a, h);
      }
    } else {
      // positions r + N2 j1 < N for j1 < N1/2; the rest are zero
      const long long base = (long long)o * in_so + plane;
      if (KIND == TK_ROWF && cg.z) {
        // p = z + beta p_old (and f += alpha p_old) at positions j1 < N1/2;
        // with lane pairs lane h updates j1 = Q4 h + i and the pair swaps
        // the new p values
        constexpr int NV = H1 / H;
        auto pos = [&](int i) { return base + (long long)(r + N2 * (NV * h + i)) * in_sk; };
        cx<T> zv[NV];
#pragma unroll
        for (int i = 0; i < NV; i++) zv[i] = cg.z[pos(i)];
        if (cg.zz_old) {
          cx<T> pv[NV];
          const cx<T> *pold = cg.pin ? cg.pin : cg.p;
#pragma unroll
          for (int i = 0; i < NV; i++) pv[i] = pold[pos(i)];
          if (cg.alpha) {
            const T al = (T)cg.alpha[plane];
            cx<T> fv[NV];
#pragma unroll
            for (int i = 0; i < NV; i++) fv[i] = cg.f[pos(i)];
#pragma unroll
            for (int i = 0; i < NV; i++) cg.f[pos(i)] = fv[i] + al * pv[i];
          }
          const T b = (T)(cg.zz[plane] / cg.zz_old[plane]);
#pragma unroll
          for (int i = 0; i < NV; i++) zv[i] = b * pv[i] + zv[i];
        }
#pragma unroll
        for (int i = 0; i < NV; i++) cg.p[pos(i)] = zv[i];
        if constexpr (H == 1) {
#pragma unroll
          for (int i = 0; i < NV; i++) a[i] = zv[i];
        } else {
#pragma unroll
          for (int i = 0; i < NV; i++) {
            cx<T> other;
            other.x = __shfl_xor(zv[i].x, 1);
            other.y = __shfl_xor(zv[i].y, 1);
            a[i] = h ? other : zv[i];
            a[NV + i] = h ? zv[i] : other;
          }
        }
      } else {
        const cx<T> *src = in + base;
#pragma unroll
        for (int j1 = 0; j1 < H1; j1++) a[j1] = src[(long long)(r + N2 * j1) * in_sk];
      }
      if constexpr (H == 1) {
        dft_half_in<T, N1, -1>(a);
      } else {
        if (h) {
#pragma unroll
          for (int j = 1; j < A; j++) a[j] = mul_wconst<T, -1, N1>(a[j], j);
        }
        dft_reg<T, A, -1>(a);
      }
    }
  }
  __syncthreads();  // twiddle table complete
  if (actA) {
#pragma unroll
    for (int s = 0; s < A; s++) {
      cx<T> w = stw[r * k1of(s)];
      if (SA > 0) w = conjx(w);
      a[s] = a[s] * w;
    }
  }
  // ---- step B (thread kb = k1 < N1) ----
  cx<T> b[N2];
  auto wA = [&](int s) { return r * rowA + k1of(s) * G + g; };
  auto rB = [&](int j2) { return j2 * rowA + kb * G + gb; };
  if constexpr (H == 1)
    xchg<T, N1, N2, SPLIT>(lx, a, actA, r * rowA + g, G, b, true, r * G + g, rowA);
  else
    xchg_at<T, A, N2, SPLIT>(lx, a, actA, wA, b, actB, rB);
  if constexpr (KIND == TK_ROWF) {
    if (actB) {
      dft_reg<T, N2, -1>(b);
      cx<T> *dst = out + (long long)o * out_so + planeB;
#pragma unroll
      for (int k2 = 0; k2 < N2; k2++) dst[(long long)(kb + N1 * k2) * out_sk] = b[k2];
    }
  } else if constexpr (KIND == TK_ROWI) {
    // outputs kb + N1 k2 < N for k2 < N2/2
    double s = 0;
    if (actB) {
      dft_half_out<T, N2, +1>(b);
      const long long ob = (long long)o * out_so + planeB;
      cx<T> *dst = out + ob;
      if (cg.add) {
        cx<T> av[H2];
#pragma unroll
        for (int k2 = 0; k2 < H2; k2++) av[k2] = cg.add[ob + (long long)(kb + N1 * k2) * out_sk];
#pragma unroll
        for (int k2 = 0; k2 < H2; k2++) b[k2] = av[k2] + b[k2];
      }
      if (cg.zu) {
        // z -= alpha q, row sum of |z|^2
        cx<T> *zp = cg.zu + ob;
        const T na = (T)(-cg.alz[planeB]);
        cx<T> zv[H2];
#pragma unroll
        for (int k2 = 0; k2 < H2; k2++) zv[k2] = zp[(long long)(kb + N1 * k2) * out_sk];
#pragma unroll
        for (int k2 = 0; k2 < H2; k2++) {
          b[k2] = zv[k2] + na * b[k2];
          s += (double)b[k2].x * b[k2].x + (double)b[k2].y * b[k2].y;
        }
        dst = zp;
      } else if (cg.p) {
        // row sum: Re <p, q> with p, |q|^2 without (when part is given)
        const cx<T> *pp = cg.p + ob;
        cx<T> pv[H2];
#pragma unroll
        for (int k2 = 0; k2 < H2; k2++) pv[k2] = pp[(long long)(kb + N1 * k2) * out_sk];
#pragma unroll
        for (int k2 = 0; k2 < H2; k2++) s += (double)pv[k2].x * b[k2].x + (double)pv[k2].y * b[k2].y;
      } else if (cg.part) {
#pragma unroll
        for (int k2 = 0; k2 < H2; k2++) s += (double)b[k2].x * b[k2].x + (double)b[k2].y * b[k2].y;
      }
#pragma unroll
      for (int k2 = 0; k2 < H2; k2++) dst[(long long)(kb + N1 * k2) * out_sk] = b[k2];
    }
    if (cg.part) {
      // fixed-order sum over the N1 step-B threads of each plane
      __syncthreads();  // exchange buffer free
      double *sh = reinterpret_cast<double *>(smem_raw);
      if (actB) sh[kb * G + gb] = s;
      __syncthreads();
      if (kb == 0) {
        double t = 0;
        for (int rr = 0; rr < N1; rr++) t += sh[rr * G + gb];
        cg.part[(long long)o * pstride + planeB] = t;
      }
    }
  } else {
    // forward N2-point DFT, real product, inverse N2-point DFT, twiddle, exchange
    const T *mrow = mid + (long long)o * n;
    T mv[N2];
    double qf = 0;  // this thread's part of sum mid3 |FFT(in)|^2 (TK_COL2) or mid |FFT(in)|^2 (TK_COL)
    if (actB) {
      dft_reg<T, N2, -1>(b);
      if constexpr (KIND == TK_COL2) {
        if (cg.mid3) {
          const T *m3 = cg.mid3 + (long long)o * n;
#pragma unroll
          for (int k2 = 0; k2 < N2; k2++) mv[k2] = m3[kb + N1 * k2];
#pragma unroll
          for (int k2 = 0; k2 < N2; k2++) qf += (double)mv[k2] * ((double)b[k2].x * b[k2].x + (double)b[k2].y * b[k2].y);
        }
      }
#pragma unroll
      for (int k2 = 0; k2 < N2; k2++) mv[k2] = mrow[kb + N1 * k2];
      if constexpr (KIND == TK_COL) {
        // <p, T p> = sum mid |P^|^2 over the spectrum (Parseval on the embedding)
        if (cg.qpart) {
#pragma unroll
          for (int k2 = 0; k2 < N2; k2++) qf += (double)mv[k2] * ((double)b[k2].x * b[k2].x + (double)b[k2].y * b[k2].y);
        }
      }
#pragma unroll
      for (int k2 = 0; k2 < N2; k2++) b[k2] = mv[k2] * b[k2];
    }
    if constexpr (KIND == TK_COL2) {
      // the second input's column through the same forward steps, times
      // mid2, subtracted
      __syncthreads();  // every thread has read the exchange buffer
      if (actA) {
        const cx<T> *src = cg.in2 + (long long)o * in_so + plane;
#pragma unroll
        for (int j1 = 0; j1 < H1; j1++) a[j1] = src[(long long)(r + N2 * j1) * in_sk];
        if constexpr (H == 1) {
          dft_half_in<T, N1, -1>(a);
        } else {
          if (h) {
#pragma unroll
            for (int j = 1; j < A; j++) a[j] = mul_wconst<T, -1, N1>(a[j], j);
          }
          dft_reg<T, A, -1>(a);
        }
#pragma unroll
        for (int s = 0; s < A; s++) a[s] = a[s] * stw[r * k1of(s)];
      }
      cx<T> b2[N2];
      if constexpr (H == 1)
        xchg<T, N1, N2, SPLIT>(lx, a, actA, r * rowA + g, G, b2, true, r * G + g, rowA);
      else
        xchg_at<T, A, N2, SPLIT>(lx, a, actA, wA, b2, actB, rB);
      if (actB) {
        dft_reg<T, N2, -1>(b2);
        const T *mrow2 = cg.mid2 + (long long)o * n;
#pragma unroll
        for (int k2 = 0; k2 < N2; k2++) mv[k2] = mrow2[kb + N1 * k2];
#pragma unroll
        for (int k2 = 0; k2 < N2; k2++) b[k2] = b[k2] - mv[k2] * b2[k2];
      }
    }
    if (actB) {
      dft_reg<T, N2, +1>(b);
#pragma unroll
      for (int m2 = 0; m2 < N2; m2++) b[m2] = b[m2] * conjx(stw[kb * m2]);
    }
    __syncthreads();  // every thread has read the exchange buffer
    // thread r = m2 gathers k1 = 0..N1-1 (lane pairs: k1 of parity h);
    // outputs r + N2 m1 < N for m1 < N1/2
    if constexpr (H == 1) {
      xchg<T, N2, N1, SPLIT>(lx, b, true, r * G + g, rowA, a, actA, r * rowA + g, G);
      if (actA) {
        dft_half_out<T, N1, +1>(a);
        cx<T> *dst = out + (long long)o * out_so + plane;
#pragma unroll
        for (int m1 = 0; m1 < H1; m1++) dst[(long long)(r + N2 * m1) * out_sk] = a[m1];
      }
    } else {
      xchg_at<T, N2, A, SPLIT>(
          lx, b, actB, [&](int m2) { return m2 * rowA + kb * G + gb; }, a, actA,
          [&](int s) { return r * rowA + (2 * s + h) * G + g; });
      if (actA) {
        dft_reg<T, A, +1>(a);
        if (h) {
#pragma unroll
          for (int m = 1; m < A; m++) a[m] = mul_wconst<T, +1, N1>(a[m], m);
        }
        pair_swap<T, Q4, false>(a, h);
        cx<T> *dst = out + (long long)o * out_so + plane;
#pragma unroll
        for (int i = 0; i < Q4; i++) dst[(long long)(r + N2 * (Q4 * h + i)) * out_sk] = a[i];
      }
    }
    if constexpr (KIND == TK_COL2 || KIND == TK_COL) {
      if (cg.qpart) {
        // fixed-order sum over the N1 step-B threads of each plane
        __syncthreads();  // exchange buffer free
        double *sh = reinterpret_cast<double *>(smem_raw);
        if (actB) sh[kb * G + gb] = qf;
        __syncthreads();
        if (kb == 0) {
          double t = 0;
          for (int rr = 0; rr < N1; rr++) t += sh[rr * G + gb];
          cg.qpart[(long long)o * pstride + planeB] = t;
        }
      }
    }
  }
}

// z-axis transforms of a [line][z] volume (z contiguous, n = N1 N2) as the
// four-step transform above with lanes along z: G = 256 / N1 lines per
// block, thread (g, r) holds inputs r + N2 j1 (r < N2) in step A and outputs
// r + N1 k2 (r < N1) in step B, so every load and store of a line is one
// contiguous run.  S(z) = (z + n/2) mod n, the reference's z shift, maps to
// a compile-time permutation of the register index: S(r + N2 j1) =
// r + N2 ((j1 + N1/2) mod N1), S^-1(k1 + N1 k2) = k1 + N1 ((k2 + N2/2) mod N2).
//   forward (INV false): d[z] = img[S(z)];  D = DFT_-(d);  fh[p] = D[S(p)];
//                        f0[z] = d[z] (when f0 is non-null)
//   inverse (INV true):  d[k] = g[S(k)] + sum_{s < nk} alpha[(s K + k') n + S(k)]
//                        pv[s pstride + S(k)] (nk = 0 when pv is null);
//                        e = DFT_+(d);  img[z] = Re e[S(z)] / n
// tw holds exp(-2 pi i e / n), e < n.  K volumes of a batch: line (y x) K + k'
// of the Fourier side is line (y x) of real-space volume k' (zimg_base).
template <typename T, int N1, int N2, bool INV>
__global__ void __launch_bounds__(256)
k_zfft2(const T *__restrict__ img_in, const cx<T> *__restrict__ g_in, const cx<T> *__restrict__ pv,
        const double *__restrict__ alpha, cx<T> *__restrict__ fh, cx<T> *__restrict__ f0, T *__restrict__ img_out,
        const cx<T> *__restrict__ tw, long long nlines, int nk, long long pstride, int K, long long vimg) {
  constexpr int n = N1 * N2, G = 256 / N1, RS = G * N2 + 1;
  constexpr int S = INV ? +1 : -1;
  static_assert(256 % N1 == 0 && N1 >= N2 && N2 % 2 == 0 && N1 % 8 == 0, "lane layout");
  __shared__ cx<T> lx[N1 * RS];
  __shared__ cx<T> stw[n];
  const int tid = threadIdx.x, r = tid % N1, g = tid / N1;
  for (int i = tid; i < n; i += 256) stw[i] = tw[i];
  const long long line = (long long)xcd_block(blockIdx.x, gridDim.x) * G + g;
  const bool act = line < nlines;
  const bool actA = act && r < N2;
  const long long base = line * n;                     // Fourier side
  const long long ibase = zimg_base(line, n, K, vimg);  // real-space side
  cx<T> a[N1];
  if (actA) {
#pragma unroll
    for (int j1 = 0; j1 < N1; j1++) {
      const int zs = r + N2 * ((j1 + N1 / 2) % N1);
      if constexpr (!INV) {
        const T v = img_in[ibase + zs];
        a[j1] = mk(v, (T)0);
      } else {
        a[j1] = g_in[base + zs];
      }
    }
    if constexpr (INV) {
      // f + sum_s alpha_s p_s, the steps in order, 8 positions at a time
      // with their loads issued together (loaded and used one by one, the
      // loads were waited for one at a time)
#pragma unroll 1
      for (int s = 0; s < nk; s++) {
        const cx<T> *ps = pv + s * pstride + base;
        // alpha_s holds one value per Fourier plane, n K of them
        const double *as = alpha + (s * (long long)K + line % K) * n;
#pragma unroll
        for (int c = 0; c < N1; c += 8) {
          cx<T> pl[8];
          double al[8];
#pragma unroll
          for (int u = 0; u < 8; u++) {
            const int zs = r + N2 * ((c + u + N1 / 2) % N1);
            pl[u] = ps[zs];
            al[u] = as[zs];
          }
#pragma unroll
          for (int u = 0; u < 8; u++) a[c + u] = a[c + u] + (T)al[u] * pl[u];
        }
      }
    } else {
      if (f0) {
#pragma unroll
        for (int j1 = 0; j1 < N1; j1++) f0[base + r + N2 * j1] = a[j1];
      }
    }
    dft_reg<T, N1, S>(a);
  }
  __syncthreads();  // twiddle table complete
  if (actA) {
#pragma unroll
    for (int k1 = 0; k1 < N1; k1++) {
      cx<T> w = stw[r * k1];
      if (S > 0) w = conjx(w);
      lx[k1 * RS + g * N2 + r] = a[k1] * w;
    }
  }
  __syncthreads();
  cx<T> b[N2];
#pragma unroll
  for (int j2 = 0; j2 < N2; j2++) b[j2] = lx[r * RS + g * N2 + j2];
  dft_reg<T, N2, S>(b);
  if (act) {
#pragma unroll
    for (int k2 = 0; k2 < N2; k2++) {
      const int p = r + N1 * ((k2 + N2 / 2) % N2);
      if constexpr (!INV) fh[base + p] = b[k2];
      else img_out[ibase + p] = b[k2].x / (T)n;
    }
  }
}

}  // namespace nudim
