// fft2.hip -- instantiations and dispatch of the fixed-length line FFTs.
#include "fft2_launch.hpp"

namespace nudim {

// threads per block of the fixed-length kernels (a tuning knob: larger
// blocks give longer contiguous plane runs, smaller ones more blocks per CU)
int g_fft2_threads = getenv("NUDIM_FFT2_THREADS") ? atoi(getenv("NUDIM_FFT2_THREADS")) : 256;


bool toep_supported(int n) {
  switch (n) {
    case 128: case 256: case 384: case 512: case 768: case 1024: case 1536: return true;
    default: return toep_supported_more(n);
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
    default:
      if (toep_launch_more<T>(kind, n, nlines, in, out, in_so, in_sk, out_so, out_sk, P, st, tw, mid, cg, pstride,
                              pairs))
        return;
      throw Error(NUDIM_EINVAL, "no fixed-length Toeplitz pass for length " + std::to_string(n));
  }
#undef NUDIM_TOEP
}

template void toep_launch<float>(int, int, int, const cx<float> *, cx<float> *, long long, long long, long long,
                                 long long, int, hipStream_t, const cx<float> *, const float *,
                                 const ToepCg<float> &, int, bool);
template void toep_launch<double>(int, int, int, const cx<double> *, cx<double> *, long long, long long, long long,
                                  long long, int, hipStream_t, const cx<double> *, const double *,
                                  const ToepCg<double> &, int, bool);

// z-axis transforms (k_zfft2) for n in {64, 96, 128, 192, 256, 512}; false when no
// fixed-length kernel exists (the caller then uses the generic kernels)
template <typename T>
bool zfft2_launch(int n, bool inv, long long nlines, const T *img_in, const cx<T> *g_in, const cx<T> *pv,
                  const double *alpha, cx<T> *fh, cx<T> *f0, T *img_out, const cx<T> *tw, hipStream_t st, int nk,
                  long long pstride, int K, long long vimg) {
#define NUDIM_ZFFT(NN, A, B)                                                                             \
  case NN:                                                                                               \
    launch_zfft2<T, A, B>(inv, nlines, img_in, g_in, pv, alpha, fh, f0, img_out, tw, st, nk, pstride, K, vimg); \
    return true;
  if (!g_use_fft2) return false;
  switch (n) {
    NUDIM_ZFFT(64, 8, 8)
    NUDIM_ZFFT(128, 16, 8)
    NUDIM_ZFFT(256, 16, 16)
    NUDIM_ZFFT(512, 32, 16)
    default:
      return zfft2_launch_more<T>(n, inv, nlines, img_in, g_in, pv, alpha, fh, f0, img_out, tw, st, nk, pstride, K,
                                  vimg);
  }
#undef NUDIM_ZFFT
}

template bool zfft2_launch<float>(int, bool, long long, const float *, const cx<float> *, const cx<float> *,
                                  const double *, cx<float> *, cx<float> *, float *, const cx<float> *, hipStream_t,
                                  int, long long, int, long long);
template bool zfft2_launch<double>(int, bool, long long, const double *, const cx<double> *, const cx<double> *,
                                   const double *, cx<double> *, cx<double> *, double *, const cx<double> *,
                                   hipStream_t, int, long long, int, long long);

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
    // (2048 = 64 x 32 would spill registers)
    default: return fft2_launch_more<T>(n, in, out, nlines, li, lo, sign, P, st, tw, mid, mid_so);
  }
}

template bool fft2_launch<float>(int, const cx<float> *, cx<float> *, int, const LineIO &, const LineIO &, int, int,
                                 hipStream_t, const cx<float> *, const cx<float> *, long long);
template bool fft2_launch<double>(int, const cx<double> *, cx<double> *, int, const LineIO &, const LineIO &, int, int,
                                  hipStream_t, const cx<double> *, const cx<double> *, long long);

}  // namespace nudim
