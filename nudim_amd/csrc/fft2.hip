// fft2.hip -- instantiations and dispatch of the fixed-length line FFTs.
#include "engine.hpp"
#include "fft2.hpp"

namespace nudim {

// threads per block of the fixed-length kernels (a tuning knob: larger
// blocks give longer contiguous plane runs, smaller ones more blocks per CU)
int g_fft2_threads = getenv("NUDIM_FFT2_THREADS") ? atoi(getenv("NUDIM_FFT2_THREADS")) : 256;
// lane pairs in the NFFT grid FFT's 48-point stage (NUDIM_TOEP_PAIRS=0: off,
// read at each launch like the Toeplitz passes' switch at each job)
static bool toep_pairs_on() {
  const char *e = getenv("NUDIM_TOEP_PAIRS");
  return !(e && std::string(e) == "0");
}

namespace {

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

}  // namespace

bool toep_supported(int n) {
  switch (n) {
    case 128: case 256: case 384: case 512: case 768: case 1024: case 1536: return true;
    default: return false;
  }
}

template <typename T>
void toep_launch(int kind, int n, int nlines, const cx<T> *in, cx<T> *out, long long in_so, long long in_sk,
                 long long out_so, long long out_sk, int P, hipStream_t st, const cx<T> *tw, const T *mid,
                 const ToepCg<T> &cg, int pstride, bool pairs) {
#define NUDIM_TOEP(NN, A, B)                                                                                   \
  case NN:                                                                                                     \
    launch_toep<T, A, B>(kind, nlines, in, out, in_so, in_sk, out_so, out_sk, P, st, tw, mid, cg, pstride, pairs); \
    return;
  switch (n) {
    NUDIM_TOEP(128, 16, 8)
    NUDIM_TOEP(256, 16, 16)
    NUDIM_TOEP(384, 24, 16)
    NUDIM_TOEP(512, 32, 16)
    NUDIM_TOEP(768, 32, 24)
    NUDIM_TOEP(1024, 32, 32)
    NUDIM_TOEP(1536, 48, 32)
    default: throw Error(NUDIM_EINVAL, "no fixed-length Toeplitz pass for length " + std::to_string(n));
  }
#undef NUDIM_TOEP
}

template void toep_launch<float>(int, int, int, const cx<float> *, cx<float> *, long long, long long, long long,
                                 long long, int, hipStream_t, const cx<float> *, const float *,
                                 const ToepCg<float> &, int, bool);
template void toep_launch<double>(int, int, int, const cx<double> *, cx<double> *, long long, long long, long long,
                                  long long, int, hipStream_t, const cx<double> *, const double *,
                                  const ToepCg<double> &, int, bool);

// z-axis transforms (k_zfft2) for n in {64, 128, 256, 512}; false when no
// fixed-length kernel exists (the caller then uses the generic kernels)
template <typename T>
bool zfft2_launch(int n, bool inv, long long nlines, const T *img_in, const cx<T> *g_in, const cx<T> *pv,
                  const double *alpha, cx<T> *fh, cx<T> *f0, T *img_out, const cx<T> *tw, hipStream_t st, int nk,
                  long long pstride) {
#define NUDIM_ZFFT(NN, A, B)                                                                                     \
  case NN: {                                                                                                     \
    const dim3 grid((unsigned)((nlines + 256 / A - 1) / (256 / A)));                                            \
    if (inv)                                                                                                     \
      hipLaunchKernelGGL((k_zfft2<T, A, B, true>), grid, dim3(256), 0, st, img_in, g_in, pv, alpha, fh, f0,      \
                         img_out, tw, nlines, nk, pstride);                                                      \
    else                                                                                                         \
      hipLaunchKernelGGL((k_zfft2<T, A, B, false>), grid, dim3(256), 0, st, img_in, g_in, pv, alpha, fh, f0,     \
                         img_out, tw, nlines, nk, pstride);                                                      \
    return true;                                                                                                 \
  }
  if (!g_use_fft2) return false;
  switch (n) {
    NUDIM_ZFFT(64, 8, 8)
    NUDIM_ZFFT(128, 16, 8)
    NUDIM_ZFFT(256, 16, 16)
    NUDIM_ZFFT(512, 32, 16)
    default: return false;
  }
#undef NUDIM_ZFFT
}

template bool zfft2_launch<float>(int, bool, long long, const float *, const cx<float> *, const cx<float> *,
                                  const double *, cx<float> *, cx<float> *, float *, const cx<float> *, hipStream_t,
                                  int, long long);
template bool zfft2_launch<double>(int, bool, long long, const double *, const cx<double> *, const cx<double> *,
                                   const double *, cx<double> *, cx<double> *, double *, const cx<double> *,
                                   hipStream_t, int, long long);

// Returns false when no fixed-length kernel exists for n (the caller then
// uses the generic Stockham kernel).
template <typename T>
bool fft2_launch(int n, const cx<T> *in, cx<T> *out, int nlines, const LineIO &li, const LineIO &lo, int sign, int P,
                 hipStream_t st, const cx<T> *tw, const cx<T> *mid, long long mid_so) {
  switch (n) {
    case 128: return launch_shape<T, 16, 8>(in, out, nlines, li, lo, sign, P, st, tw, mid, mid_so);
    case 256: return launch_shape<T, 16, 16>(in, out, nlines, li, lo, sign, P, st, tw, mid, mid_so);
    case 384: return launch_shape<T, 24, 16>(in, out, nlines, li, lo, sign, P, st, tw, mid, mid_so);
    case 512: return launch_shape<T, 32, 16>(in, out, nlines, li, lo, sign, P, st, tw, mid, mid_so);
    case 768: return launch_shape<T, 32, 24>(in, out, nlines, li, lo, sign, P, st, tw, mid, mid_so);
    case 1024: return launch_shape<T, 32, 32>(in, out, nlines, li, lo, sign, P, st, tw, mid, mid_so);
    case 1536: return launch_shape<T, 48, 32>(in, out, nlines, li, lo, sign, P, st, tw, mid, mid_so);
    default: return false;  // (2048 = 64 x 32 would spill registers)
  }
}

template bool fft2_launch<float>(int, const cx<float> *, cx<float> *, int, const LineIO &, const LineIO &, int, int,
                                 hipStream_t, const cx<float> *, const cx<float> *, long long);
template bool fft2_launch<double>(int, const cx<double> *, cx<double> *, int, const LineIO &, const LineIO &, int, int,
                                  hipStream_t, const cx<double> *, const cx<double> *, long long);

}  // namespace nudim
