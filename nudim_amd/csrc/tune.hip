// tune.hip -- choice of the FP32 NFFT grid size and window width.
//
// The reference fixes its NFFT at n = 2 nextpow2(np) with 14 Kaiser-Bessel
// taps in FP64 (src/reconstruction.c:393-405).  The FP32 form needs far less:
// its NFFT feeds the known-node adjoints at load, the initial INFFT and,
// when R is taken through the NFFT, the R sums, where a relative error near
// FP32 rounding is enough.  For a job's
// node set this measures, on the device, candidate (n, W) pairs:
//   - n among the fixed-length FFT sizes from N + 16 up to the default
//     nice_fft_size(2N) (so a job never holds more than the default sizing
//     promises to the admission estimate);
//   - for each n the smallest W whose trafo and adjoint, on 64 random planes,
//     are within kTuneTol relative L2 of the FP64 NFFT at the reference's
//     sizes;
//   - the device time of one trafo + adjoint of those 64 planes;
// and returns the fastest candidate within the tolerance.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <random>
#include <vector>

#include "engine.hpp"
#include "geom.hpp"

namespace nudim {

namespace {

constexpr double kTuneTol = 2e-6;
constexpr int kTunePlanes = 64;
const int kTuneW[] = {4, 6, 8, 10, 12};
const int kFixedN[] = {128, 256, 384, 512, 768, 1024, 1536};

template <typename T>
void to_dev(DevBuf &b, const std::vector<double> &h) {
  std::vector<cx<T>> v(h.size() / 2);
  for (size_t i = 0; i < v.size(); i++) v[i] = mk((T)h[2 * i], (T)h[2 * i + 1]);
  b.alloc(v.size() * sizeof(cx<T>));
  NUDIM_HIP(hipMemcpy(b.p, v.data(), v.size() * sizeof(cx<T>), hipMemcpyHostToDevice));
}

template <typename T>
std::vector<double> to_host(const DevBuf &b, size_t count) {
  std::vector<cx<T>> v(count);
  NUDIM_HIP(hipMemcpy(v.data(), b.p, count * sizeof(cx<T>), hipMemcpyDeviceToHost));
  std::vector<double> h(2 * count);
  for (size_t i = 0; i < count; i++) {
    h[2 * i] = (double)v[i].x;
    h[2 * i + 1] = (double)v[i].y;
  }
  return h;
}

double rel_l2(const std::vector<double> &a, const std::vector<double> &b) {
  double num = 0, den = 0;
  for (size_t i = 0; i < a.size(); i++) {
    num += (a[i] - b[i]) * (a[i] - b[i]);
    den += b[i] * b[i];
  }
  return den > 0 ? std::sqrt(num / den) : std::sqrt(num);
}

}  // namespace

NfftTuning tune_nfft_fp32(int N, const double *x, int M, hipStream_t st, std::vector<NfftTuning> *trace) {
  const int P = kTunePlanes;
  const size_t NN = (size_t)N * N * P, MM = (size_t)M * P;
  std::mt19937_64 rng(12345);
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  std::vector<double> fhat(2 * NN), f(2 * MM);
  for (auto &v : fhat) v = u(rng);
  for (auto &v : f) v = u(rng);
  // FP64 NFFT at the reference's sizes
  std::vector<double> ref_f, ref_fhat;
  {
    NfftOp<double> r;
    r.init(N, 2 * next_pow2(N), 14, x, M, P);
    DevBuf a, b;
    to_dev<double>(a, fhat);
    b.alloc(MM * sizeof(cx<double>));
    r.trafo(a.as<cx<double>>(), b.as<cx<double>>(), st);
    NUDIM_HIP(hipStreamSynchronize(st));
    ref_f = to_host<double>(b, MM);
    to_dev<double>(a, f);
    b.alloc(NN * sizeof(cx<double>));
    r.adjoint(a.as<cx<double>>(), b.as<cx<double>>(), st);
    NUDIM_HIP(hipStreamSynchronize(st));
    ref_fhat = to_host<double>(b, NN);
  }
  DevBuf dfhat, df, out_m, out_n;
  to_dev<float>(dfhat, fhat);
  to_dev<float>(df, f);
  out_m.alloc(MM * sizeof(cx<float>));
  out_n.alloc(NN * sizeof(cx<float>));
  hipEvent_t e0, e1;
  NUDIM_HIP(hipEventCreate(&e0));
  NUDIM_HIP(hipEventCreate(&e1));
  const int nmax = nice_fft_size(2 * N);
  NfftTuning best{nmax, 8, -1.0, -1.0};
  try {
    for (int n : kFixedN) {
      if (n < N + 16 || n > nmax) continue;
      for (int W : kTuneW) {
        if (W > n / 4) break;
        NfftOp<float> op;
        op.init(N, n, W, x, M, P);
        op.trafo(dfhat.as<cx<float>>(), out_m.as<cx<float>>(), st);
        op.adjoint(df.as<cx<float>>(), out_n.as<cx<float>>(), st);
        NUDIM_HIP(hipStreamSynchronize(st));
        const double err = std::max(rel_l2(to_host<float>(out_m, MM), ref_f), rel_l2(to_host<float>(out_n, NN), ref_fhat));
        NfftTuning c{n, W, err, -1.0};
        if (err <= kTuneTol) {
          float best_ms = 1e30f;
          for (int rep = 0; rep < 3; rep++) {
            NUDIM_HIP(hipEventRecord(e0, st));
            op.trafo(dfhat.as<cx<float>>(), out_m.as<cx<float>>(), st);
            op.adjoint(df.as<cx<float>>(), out_n.as<cx<float>>(), st);
            NUDIM_HIP(hipEventRecord(e1, st));
            NUDIM_HIP(hipEventSynchronize(e1));
            float ms = 0;
            NUDIM_HIP(hipEventElapsedTime(&ms, e0, e1));
            best_ms = std::min(best_ms, ms);
          }
          c.ms = best_ms;
        }
        if (trace) trace->push_back(c);
        if (c.ms >= 0) {
          if (best.ms < 0 || c.ms < best.ms) best = c;
          break;  // a wider window at this n only costs more
        }
      }
    }
  } catch (...) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return best;
}

}  // namespace nudim
