// fft2_launch.hpp -- launch shapes of the fixed-length line FFTs and
// Toeplitz passes (fft2.hpp), shared by the translation units that
// instantiate them for different lengths (fft2.hip, fft2_more.hip).
#pragma once

#include <string>

#include "engine.hpp"
#include "fft2.hpp"

namespace nudim {

// threads per block of the fixed-length kernels (fft2.hip)
extern int g_fft2_threads;

// lane pairs in the NFFT grid FFT's 48-point stage (NUDIM_TOEP_PAIRS=0: off,
// read at each launch like the Toeplitz passes' switch at each job)
inline bool toep_pairs_on() {
  const char *e = getenv("NUDIM_TOEP_PAIRS");
  return !(e && std::string(e) == "0");
}

// FP64 generic-form lines on lane pairs (NUDIM_FFT64_PAIRS=0: the one-lane kernel)
inline bool fft2_pairs64_on() {
  static const bool on = !(getenv("NUDIM_FFT64_PAIRS") && std::string(getenv("NUDIM_FFT64_PAIRS")) == "0");
  return on;
}

template <typename T, int N1, int N2>
bool launch_shape(const cx<T> *in, cx<T> *out, int nlines, const LineIO &li, const LineIO &lo, int sign, int P,
                  hipStream_t st, const cx<T> *tw, const cx<T> *mid, long long mid_so) {
  constexpr int n = N1 * N2;
  if (mid) return false;  // the pointwise-product form runs on the Toeplitz passes
  // n/2 elements centred on position 0: the NFFT's N = n/2 modes
  const bool hin = li.cnt * 2 == n && li.shift == -n / 4;
  const bool hout = !hin && lo.cnt * 2 == n && lo.shift == -n / 4;
  // the half-input 48-point stage runs on lane pairs in 512-thread blocks
  // (as the Toeplitz row passes of n = 1536)
  const bool pairs = N1 == 48 && hin && toep_pairs_on();
  if constexpr (sizeof(T) == 8 && N1 >= 24 && N1 % 2 == 0 && N2 % 2 == 0) {
    // FP64 lines of the generic form on lane pairs (k_fft2p)
    if (!hin && !hout && fft2_pairs64_on()) {
      constexpr int TL2 = 2 * (N1 > N2 ? N1 : N2);
      // 256-thread blocks (4 planes at N1 = N2 = 32), three per CU
      int gs = 0;
      while ((TL2 << (gs + 1)) <= 256 && P % (2 << gs) == 0) gs++;
      auto lds2 = [&](int q) { return (size_t)N2 * (N1 * (1 << q) + (1 << q)) * sizeof(T) + (size_t)n * sizeof(cx<T>); };
      while (gs > 0 && lds2(gs) > 160 * 1024) gs--;
      const int nch = P >> gs;
      const dim3 gr2((unsigned)((long long)nlines * nch)), bl2(TL2 << gs);
      if (sign < 0)
        hipLaunchKernelGGL((k_fft2p<T, N1, N2, -1>), gr2, bl2, lds2(gs), st, in, out, li, lo, tw, gs, nch);
      else
        hipLaunchKernelGGL((k_fft2p<T, N1, N2, +1>), gr2, bl2, lds2(gs), st, in, out, li, lo, tw, gs, nch);
      return true;
    }
  }
  const int TL = pairs ? std::max(N1, 2 * N2) : N1, tmax = pairs ? 512 : 256;
  int cap = std::max(1, std::min(tmax, g_fft2_threads * tmax / 256) / TL);
  int Gs = 0;
  while ((2 << Gs) <= cap && P % (2 << Gs) == 0) Gs++;
  auto lds_of = [&](int gs) { return ((size_t)N2 * (N1 * (1 << gs) + (1 << gs)) + n) * sizeof(cx<T>); };
  while (Gs > 0 && lds_of(Gs) > 160 * 1024) Gs--;
  const size_t lds = lds_of(Gs);
  if (lds > 160 * 1024) return false;
  const int G = 1 << Gs;
  const int nchunk = P >> Gs;
  dim3 grid((unsigned)((long long)nlines * nchunk)), block(TL * G);
#define NUDIM_FFT2(SG, HI, HO, HH) \
  hipLaunchKernelGGL((k_fft2<T, N1, N2, SG, HI, HO, HH>), grid, block, lds, st, in, out, li, lo, tw, Gs, nchunk)
  if constexpr (N1 == 48) {
    if (pairs) {
      if (sign < 0) NUDIM_FFT2(-1, true, false, 2);
      else NUDIM_FFT2(+1, true, false, 2);
      return true;
    }
  }
  if (sign < 0) {
    if (hin) NUDIM_FFT2(-1, true, false, 1);
    else if (hout) NUDIM_FFT2(-1, false, true, 1);
    else NUDIM_FFT2(-1, false, false, 1);
  } else {
    if (hin) NUDIM_FFT2(+1, true, false, 1);
    else if (hout) NUDIM_FFT2(+1, false, true, 1);
    else NUDIM_FFT2(+1, false, false, 1);
  }
#undef NUDIM_FFT2
  return true;
}

constexpr int ilog2_floor(int v) { return v > 1 ? 1 + ilog2_floor(v / 2) : 0; }

template <typename T, int N1, int N2, int H>
void launch_toep_h(int kind, int nlines, const cx<T> *in, cx<T> *out, long long in_so, long long in_sk,
                   long long out_so, long long out_sk, int P, hipStream_t st, const cx<T> *tw, const T *mid,
                   const ToepCg<T> &cg, int pstride) {
  constexpr int n = N1 * N2, TL = (N1 > H * N2 ? N1 : H * N2);  // threads per plane
  // the row passes run 512-thread blocks: twice the planes per block, so
  // twice the contiguous bytes per line position (C3: row forward 383 -> 371
  // us with 107 KB of LDS, row inverse 242 -> 226 us; 512^3 lane pairs: 953
  // -> 849 us and 631 -> 509 us); the column passes lose more occupancy
  // than they gain
  const int tmax = (kind == TK_ROWI || kind == TK_ROWF) ? 512 : 256;
  int cap = std::max(1, std::min(tmax, g_fft2_threads * tmax / 256) / TL);
  int Gs = 0;
  while ((2 << Gs) <= cap && P % (2 << Gs) == 0) Gs++;
  auto lds_of = [&](int gs) {
    return (size_t)N2 * (N1 * (1 << gs) + (1 << gs)) * (toep_split(kind) ? 1 : 2) * sizeof(T) +
           (size_t)n * sizeof(cx<T>);
  };
  while (Gs > 0 && lds_of(Gs) > 160 * 1024) Gs--;  // a block's LDS within 160 KiB
  const int G = 1 << Gs;
  const size_t lds = lds_of(Gs);
  const int nchunk = P >> Gs;
  dim3 grid((unsigned)((long long)nlines * nchunk)), block(TL * G);
  // at the full plane group (P a multiple of it) G is a compile-time
  // constant: the exchanges' LDS addresses become constant offsets (C3 column
  // pass 372 -> 358 us, two-input column pass 820 -> 763 us)
  constexpr int GC = ilog2_floor(256 / TL), GR = ilog2_floor(512 / TL);
#define NUDIM_K(KIND, GSC)                                                                                      \
  hipLaunchKernelGGL((k_toep<T, N1, N2, KIND, H, GSC>), grid, block, lds, st, in, out, in_so, in_sk, out_so, out_sk, \
                     tw, Gs, nchunk, mid, cg, pstride)
  switch (kind) {
    case TK_ROWF:
      if (Gs == GR) NUDIM_K(TK_ROWF, GR);
      else NUDIM_K(TK_ROWF, -1);
      break;
    case TK_COL:
      if (Gs == GC) NUDIM_K(TK_COL, GC);
      else NUDIM_K(TK_COL, -1);
      break;
    case TK_COL2:
      if (Gs == GC) NUDIM_K(TK_COL2, GC);
      else NUDIM_K(TK_COL2, -1);
      break;
    default:
      if (Gs == GR) NUDIM_K(TK_ROWI, GR);
      else NUDIM_K(TK_ROWI, -1);
  }
#undef NUDIM_K
}

// lane pairs (H = 2) for the 48-point stage of n = 1536 unless pairs is false
template <typename T, int N1, int N2>
void launch_toep(int kind, int nlines, const cx<T> *in, cx<T> *out, long long in_so, long long in_sk,
                 long long out_so, long long out_sk, int P, hipStream_t st, const cx<T> *tw, const T *mid,
                 const ToepCg<T> &cg, int pstride, bool pairs) {
  if constexpr (N1 == 48) {
    if (pairs) {
      launch_toep_h<T, N1, N2, 2>(kind, nlines, in, out, in_so, in_sk, out_so, out_sk, P, st, tw, mid, cg, pstride);
      return;
    }
  }
  launch_toep_h<T, N1, N2, 1>(kind, nlines, in, out, in_so, in_sk, out_so, out_sk, P, st, tw, mid, cg, pstride);
}


// z transforms of [line][z] volumes on k_zfft2<T, A, B>
template <typename T, int A, int B>
void launch_zfft2(bool inv, long long nlines, const T *img_in, const cx<T> *g_in, const cx<T> *pv,
                  const double *alpha, cx<T> *fh, cx<T> *f0, T *img_out, const cx<T> *tw, hipStream_t st, int nk,
                  long long pstride, int K, long long vimg) {
  const dim3 grid((unsigned)((nlines + 256 / A - 1) / (256 / A)));
  if (inv)
    hipLaunchKernelGGL((k_zfft2<T, A, B, true>), grid, dim3(256), 0, st, img_in, g_in, pv, alpha, fh, f0, img_out, tw,
                       nlines, nk, pstride, K, vimg);
  else
    hipLaunchKernelGGL((k_zfft2<T, A, B, false>), grid, dim3(256), 0, st, img_in, g_in, pv, alpha, fh, f0, img_out,
                       tw, nlines, nk, pstride, K, vimg);
}

// the lengths instantiated in fft2_more.hip (line FFT / Toeplitz lengths
// 160, 192, 240, 288, 320, 480, 576, 640; z lengths 96, 192); false when n
// is not one of them
template <typename T>
bool fft2_launch_more(int n, const cx<T> *in, cx<T> *out, int nlines, const LineIO &li, const LineIO &lo, int sign,
                      int P, hipStream_t st, const cx<T> *tw, const cx<T> *mid, long long mid_so);
template <typename T>
bool toep_launch_more(int kind, int n, int nlines, const cx<T> *in, cx<T> *out, long long in_so, long long in_sk,
                      long long out_so, long long out_sk, int P, hipStream_t st, const cx<T> *tw, const T *mid,
                      const ToepCg<T> &cg, int pstride, bool pairs);
template <typename T>
bool zfft2_launch_more(int n, bool inv, long long nlines, const T *img_in, const cx<T> *g_in, const cx<T> *pv,
                       const double *alpha, cx<T> *fh, cx<T> *f0, T *img_out, const cx<T> *tw, hipStream_t st,
                       int nk, long long pstride, int K, long long vimg);
bool toep_supported_more(int n);

}  // namespace nudim
