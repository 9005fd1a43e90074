// fft2_more.hip -- the fixed-length line FFTs, Toeplitz passes and z
// transforms for the lengths of smaller and uneven volumes: line and
// Toeplitz lengths 160, 192, 240, 288, 320, 480, 576 and 640 (2 np for np =
// 80 ... 320), z lengths 96 and 192.  A translation unit of its own so that
// it compiles beside fft2.hip.
#include "fft2_launch.hpp"

namespace nudim {

bool toep_supported_more(int n) {
  switch (n) {
    case 160: case 192: case 240: case 288: case 320: case 480: case 576: case 640: return true;
    default: return false;
  }
}

// n = N1 N2 with N1 >= N2, both even; N1 and N2 multiples of 4 where the
// NFFT's centred half-length input or output needs them
template <typename T>
bool fft2_launch_more(int n, const cx<T> *in, cx<T> *out, int nlines, const LineIO &li, const LineIO &lo, int sign,
                      int P, hipStream_t st, const cx<T> *tw, const cx<T> *mid, long long mid_so) {
  switch (n) {
    case 160: return launch_shape<T, 20, 8>(in, out, nlines, li, lo, sign, P, st, tw, mid, mid_so);
    case 192: return launch_shape<T, 16, 12>(in, out, nlines, li, lo, sign, P, st, tw, mid, mid_so);
    case 240: return launch_shape<T, 20, 12>(in, out, nlines, li, lo, sign, P, st, tw, mid, mid_so);
    case 288: return launch_shape<T, 24, 12>(in, out, nlines, li, lo, sign, P, st, tw, mid, mid_so);
    case 320: return launch_shape<T, 20, 16>(in, out, nlines, li, lo, sign, P, st, tw, mid, mid_so);
    case 480: return launch_shape<T, 24, 20>(in, out, nlines, li, lo, sign, P, st, tw, mid, mid_so);
    case 576: return launch_shape<T, 24, 24>(in, out, nlines, li, lo, sign, P, st, tw, mid, mid_so);
    case 640: return launch_shape<T, 32, 20>(in, out, nlines, li, lo, sign, P, st, tw, mid, mid_so);
    default: return false;
  }
}

template <typename T>
bool toep_launch_more(int kind, int n, int nlines, const cx<T> *in, cx<T> *out, long long in_so, long long in_sk,
                      long long out_so, long long out_sk, int P, hipStream_t st, const cx<T> *tw, const T *mid,
                      const ToepCg<T> &cg, int pstride, bool pairs) {
#define NUDIM_TOEP(NN, A, B)                                                                                      \
  case NN:                                                                                                        \
    launch_toep<T, A, B>(kind, nlines, in, out, in_so, in_sk, out_so, out_sk, P, st, tw, mid, cg, pstride, pairs); \
    return true;
  switch (n) {
    NUDIM_TOEP(160, 20, 8)
    NUDIM_TOEP(192, 16, 12)
    NUDIM_TOEP(240, 20, 12)
    NUDIM_TOEP(288, 24, 12)
    NUDIM_TOEP(320, 20, 16)
    NUDIM_TOEP(480, 24, 20)
    NUDIM_TOEP(576, 24, 24)
    NUDIM_TOEP(640, 32, 20)
    default: return false;
  }
#undef NUDIM_TOEP
}

template <typename T>
bool zfft2_launch_more(int n, bool inv, long long nlines, const T *img_in, const cx<T> *g_in, const cx<T> *pv,
                       const double *alpha, cx<T> *fh, cx<T> *f0, T *img_out, const cx<T> *tw, hipStream_t st,
                       int nk, long long pstride, int K, long long vimg) {
  switch (n) {
    case 96:
      launch_zfft2<T, 16, 6>(inv, nlines, img_in, g_in, pv, alpha, fh, f0, img_out, tw, st, nk, pstride, K, vimg);
      return true;
    case 192:
      launch_zfft2<T, 16, 12>(inv, nlines, img_in, g_in, pv, alpha, fh, f0, img_out, tw, st, nk, pstride, K, vimg);
      return true;
    default: return false;
  }
}

#define NUDIM_MORE(T)                                                                                              \
  template bool fft2_launch_more<T>(int, const cx<T> *, cx<T> *, int, const LineIO &, const LineIO &, int, int,    \
                                    hipStream_t, const cx<T> *, const cx<T> *, long long);                         \
  template bool toep_launch_more<T>(int, int, int, const cx<T> *, cx<T> *, long long, long long, long long,        \
                                    long long, int, hipStream_t, const cx<T> *, const T *, const ToepCg<T> &, int, \
                                    bool);                                                                         \
  template bool zfft2_launch_more<T>(int, bool, long long, const T *, const cx<T> *, const cx<T> *, const double *, \
                                     cx<T> *, cx<T> *, T *, const cx<T> *, hipStream_t, int, long long, int,         \
                                     long long);
NUDIM_MORE(float)
NUDIM_MORE(double)
#undef NUDIM_MORE

}  // namespace nudim
