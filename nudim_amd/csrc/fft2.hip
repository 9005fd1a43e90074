// fft2.hip -- instantiations and dispatch of the fixed-length line FFTs.
#include "engine.hpp"
#include "fft2.hpp"

namespace nudim {

// threads per block of the fixed-length kernels (a tuning knob: larger
// blocks give longer contiguous plane runs, smaller ones more blocks per CU)
int g_fft2_threads = getenv("NUDIM_FFT2_THREADS") ? atoi(getenv("NUDIM_FFT2_THREADS")) : 256;

namespace {

template <typename T, int N1, int N2>
bool launch_shape(const cx<T> *in, cx<T> *out, int nlines, const LineIO &li, const LineIO &lo, int sign, int P,
                  hipStream_t st, const cx<T> *tw, const cx<T> *mid, long long mid_so) {
  constexpr int n = N1 * N2;
  int cap = std::max(1, std::min(256, g_fft2_threads) / N1);
  int Gs = 0;
  while ((2 << Gs) <= cap && P % (2 << Gs) == 0) Gs++;
  const int G = 1 << Gs;
  size_t lds = ((size_t)N2 * (N1 * G + G) + n) * sizeof(cx<T>);
  if (lds > 160 * 1024) return false;
  const int nchunk = P >> Gs;
  dim3 grid((unsigned)((long long)nlines * nchunk)), block(N1 * G);
  if (mid) {
    hipLaunchKernelGGL((k_fft2<T, N1, N2, -1, 1>), grid, block, lds, st, in, out, li, lo, tw, Gs, nchunk, mid, mid_so);
  } else if (sign < 0) {
    hipLaunchKernelGGL((k_fft2<T, N1, N2, -1, 0>), grid, block, lds, st, in, out, li, lo, tw, Gs, nchunk, mid, mid_so);
  } else {
    hipLaunchKernelGGL((k_fft2<T, N1, N2, +1, 0>), grid, block, lds, st, in, out, li, lo, tw, Gs, nchunk, mid, mid_so);
  }
  return true;
}

}  // namespace

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
