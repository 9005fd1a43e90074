// api.hip -- reconstruction job context and the C ABI of libnudim_hip.so.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "engine.hpp"
#include "spread.hpp"
#include "nudim.h"

using namespace nudim;

static thread_local std::string g_last_error;

static int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

#define API_BEGIN try {
#define API_END                                            \
  }                                                        \
  catch (const nudim::Error &e) { return fail(e.code, e.what()); } \
  catch (const std::bad_alloc &) { return fail(NUDIM_ENOMEM, "host allocation failed"); } \
  catch (const std::exception &e) { return fail(NUDIM_EINVAL, e.what()); }

extern "C" const char *nudim_last_error(void) { return g_last_error.c_str(); }

void nudim::set_last_error(const std::string &msg) { g_last_error = msg; }

// ---------------------------------------------------------------------------
// device kernels of the set-up stage
// ---------------------------------------------------------------------------

// projection FT, reconstruction.c:306-367, as the separable exact sum:
// Q[t][r][i] = sum_c P[t][r][c] ex[c][i]
__global__ void k_pft_x(const double *__restrict__ proj, const cx<double> *__restrict__ ex, cx<double> *__restrict__ Q,
                        int nx, int np, long long rows) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * np) return;
  long long row = i / np;
  int k = (int)(i % np);
  const double *pr = proj + row * nx;
  cx<double> acc = mk(0.0, 0.0);
  for (int c = 0; c < nx; c++) {
    double v = pr[c];
    if (v == 0.0) continue;
    cx<double> e = ex[(long long)c * np + k];
    acc = acc + mk(v * e.x, v * e.y);
  }
  Q[i] = acc;
}

// tomo[t][j][i] = sum_r Q[t][r][i] ez[r][j]
__global__ void k_pft_z(const cx<double> *__restrict__ Q, const cx<double> *__restrict__ ez, cx<double> *__restrict__ tomo,
                        int nz, int np, int nmeas) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long total = (long long)nmeas * nz * np;
  if (i >= total) return;
  int k = (int)(i % np);
  long long tj = i / np;
  int j = (int)(tj % nz);
  long long t = tj / nz;
  const cx<double> *q = Q + t * nz * np + k;
  cx<double> acc = mk(0.0, 0.0);
  for (int r = 0; r < nz; r++) acc = acc + q[(long long)r * np] * ez[(long long)r * nz + j];
  tomo[i] = acc;
}

// sample assembly, reconstruction.c:437-484: meas[m][off + j] of a sample
// table with Pt planes per node (off = k nz for volume k of a batch)
template <typename T>
__global__ void k_assemble(const cx<double> *__restrict__ tomo, const int *__restrict__ kind, const int *__restrict__ order,
                           const double *__restrict__ rnd, cx<T> *__restrict__ meas, int M, int np, int nz, int Pt,
                           int off) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)M * nz) return;
  int m = (int)(i / nz), j = (int)(i % nz);
  cx<T> v = mk((T)0, (T)0);
  if (kind[m] == 1) {
    cx<double> t = tomo[((long long)order[m] * nz + j) * np + (m % np)];
    v = mk((T)t.x, (T)t.y);
  } else if (rnd) {
    v = mk((T)rnd[(long long)j * M + m], (T)0);
  }
  meas[(long long)m * Pt + off + j] = v;
}

// R sums from the quadratic form (reconstruction.c:574-582 over the known
// nodes K): sum_K |A f - m|^2 = f^H T_K f - 2 Re <f, A_K^H m> + sum_K |m|^2,
// and the same with + for |A f + m|^2.  Near a consistent f the first sum is
// a cancellation of terms of size |m|^2, each carrying the FP32 transform's
// relative error (~1e-6), so it can come out slightly negative: both sums
// are clamped at zero (R's absolute accuracy floor is then about 1e-3).
__global__ void k_rform_final(const double *__restrict__ q, const double *__restrict__ x,
                              const double *__restrict__ m2, double *__restrict__ d1, double *__restrict__ d2, int P) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  d1[p] = fmax(q[p] - 2.0 * x[p] + m2[p], 0.0);
  d2[p] = fmax(q[p] + 2.0 * x[p] + m2[p], 0.0);
}

// R = sqrt(sum_p d1 / sum_p d2) (reconstruction.c:603-604); one block of 256
// threads per volume (block k: planes k P .. k P + P - 1 into out[k]),
// fixed-order sums
__global__ void k_rfactor_final(const double *__restrict__ d1, const double *__restrict__ d2, int P,
                                double *__restrict__ out) {
  __shared__ double sa[256], sb[256];
  d1 += (long long)blockIdx.x * P;
  d2 += (long long)blockIdx.x * P;
  out += blockIdx.x;
  double a = 0, b = 0;
  for (int p = threadIdx.x; p < P; p += 256) {
    a += d1[p];
    b += d2[p];
  }
  sa[threadIdx.x] = a;
  sb[threadIdx.x] = b;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) {
      sa[threadIdx.x] += sa[threadIdx.x + h];
      sb[threadIdx.x] += sb[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sqrt(sa[0] / sb[0]);
}

// ---------------------------------------------------------------------------
// projection FT on the device (FP64)
// ---------------------------------------------------------------------------

static void projection_ft_dev(int nx, int nz, int np, int nmeas, const double *dproj, cx<double> *dtomo,
                              hipStream_t st) {
  std::vector<double> x;
  uniform_nodes(np, nz, &x);
  std::vector<double> ex(2 * (size_t)nx * np), ez(2 * (size_t)nz * nz);
  for (int c = 0; c < nx; c++)
    for (int i = 0; i < np; i++) {
      double a = -2.0 * M_PI * (double)(c - nx / 2) * x[2 * i + 0];
      ex[2 * ((size_t)c * np + i)] = std::cos(a);
      ex[2 * ((size_t)c * np + i) + 1] = std::sin(a);
    }
  for (int r = 0; r < nz; r++)
    for (int j = 0; j < nz; j++) {
      double a = -2.0 * M_PI * (double)(r - nz / 2) * x[2 * ((size_t)j * np) + 1];
      ez[2 * ((size_t)r * nz + j)] = std::cos(a);
      ez[2 * ((size_t)r * nz + j) + 1] = std::sin(a);
    }
  DevBuf dex(ex.size() * 8), dez(ez.size() * 8), dQ((size_t)nmeas * nz * np * sizeof(cx<double>));
  NUDIM_HIP(hipMemcpyAsync(dex.p, ex.data(), ex.size() * 8, hipMemcpyHostToDevice, st));
  NUDIM_HIP(hipMemcpyAsync(dez.p, ez.data(), ez.size() * 8, hipMemcpyHostToDevice, st));
  long long rows = (long long)nmeas * nz;
  hipLaunchKernelGGL(k_pft_x, dim3(cdiv(rows * np, 256)), dim3(256), 0, st, dproj, dex.as<cx<double>>(),
                     dQ.as<cx<double>>(), nx, np, rows);
  hipLaunchKernelGGL(k_pft_z, dim3(cdiv(rows * np, 256)), dim3(256), 0, st, dQ.as<cx<double>>(), dez.as<cx<double>>(),
                     dtomo, nz, np, nmeas);
  NUDIM_HIP(hipGetLastError());
  NUDIM_HIP(hipStreamSynchronize(st));
}

// ---------------------------------------------------------------------------
// job
// ---------------------------------------------------------------------------

// A job reconstructs nvol volumes of one geometry together (a batch; 1 for
// a plain job): volume k's Fourier planes are k nz .. k nz + nz - 1 of every
// Fourier-side operator, its real-space volume the k-th of nvol consecutive
// [y][x][z] volumes.  Every operator works plane by plane, so each volume
// follows the same arithmetic as its own job.
struct JobBase {
  virtual ~JobBase() = default;
  virtual void load_meas_planes(const double *slices) = 0;  // host c128 [nz][M] (nvol = 1)
  // host projections of volume vol; the known nodes' adjoint terms are formed
  // once every volume is loaded (finish_load)
  virtual void load_projections(const double *proj, const double *rnd, int vol) = 0;
  virtual void finish_load() = 0;
  virtual void initial() = 0;
  // one pass; R of volume k into rdev_slot[k]
  virtual void iterate(int nn, double *rdev_slot) = 0;
  virtual void get_volume(double *img, int vol) = 0;
  virtual void set_volume(const double *img) = 0;  // nvol = 1
  virtual size_t bytes() const = 0;
  // the host-side state a pass reads and changes: which buffer holds the
  // current, the previous and the spare volume
  virtual void get_state(void *s[3]) const = 0;
  virtual void set_state(void *const s[3]) = 0;
  double *dump_before = nullptr, *dump_after = nullptr;
  int dump_nn = -1;
};

struct Geometry {
  int nmeas = 0, nang = 0, M = 0;
  double pad_min = 0, pad_max = 0;
  std::vector<double> x, w;
  std::vector<int> kind, order;
};

static void validate(const nudim_params *p) {
  NUDIM_CHECK(p, "null parameters");
  NUDIM_CHECK(p->nx > 0 && p->nz > 0, "You need to provide valid projection size with --nx, --nz.");
  NUDIM_CHECK(p->np > 0, "You need to provide valid projection dimension in XY plane with --pad.");
  NUDIM_CHECK(p->nhio >= 0, "You need to provide valid nhio with --nhio.");
  NUDIM_CHECK(p->ner >= 0, "You need to provide valid ner with --ner.");
  NUDIM_CHECK(p->interval > 0, "You need to provide valid angular sampling interval with --interval.");
  NUDIM_CHECK(!(p->padangle < 0 ||
                p->interval / p->padangle - (int)(p->interval / p->padangle + 0.1) > 1e-5) &&
                  p->padangle > 0,
              "You need to provide valid padded angular sampling interval with --padangle.");
  NUDIM_CHECK(!(p->gam_max < -90 || p->gam_max > 90 || p->gam_min < -90),
              "You need to provide valid scope of tomogram angles (-90,90] with --max, --min.");
  NUDIM_CHECK(!(p->config_height && (p->nheight <= 0 || p->ncenter <= 0)),
              "You must provide valid both sample height and center with --height and --center.");
  NUDIM_CHECK(p->np % 2 == 0 && p->nz % 2 == 0, "np and nz must be even on this path");
  NUDIM_CHECK(p->nx <= p->np, "--nx must not exceed --pad");
  NUDIM_CHECK(p->precision == NUDIM_FP64_PLAIN || p->precision == NUDIM_FP32, "unknown precision");
  NUDIM_CHECK(!(p->denoise && p->method == NUDIM_DIFFUSION && p->nz != p->np),
              "DIFFUSION denoising needs nz == np (the reference clamps x with the z extent)");
  // the TV sweep kernels address the volume with 32-bit byte offsets
  NUDIM_CHECK(!(p->denoise && p->method == NUDIM_TV &&
                (double)p->np * p->np * p->nz * (p->precision == NUDIM_FP32 ? 4 : 8) >= 4294967296.0),
              "TV denoising needs a volume of fewer than 2^32 bytes (np^2 nz x 4 in FP32, x 8 in FP64)");
}

static void make_geometry(const nudim_params *p, Geometry *g) {
  g->nmeas = (int)((p->gam_max - p->gam_min) / p->interval + 0.5) + 1;
  pad_range(p->gam_min, p->interval, &g->pad_min, &g->pad_max);
  g->M = polar_nodes(g->pad_min, g->pad_max, p->padangle, p->np, &g->x);
  g->nang = g->M / p->np;
  if (p->weight) density_weights(p->np, g->nang, p->padangle, &g->w);
  sample_classes(p->np, g->nang, g->pad_min, p->padangle, p->gam_min, p->gam_max, p->interval, &g->kind,
                 &g->order);
}

// Tuned FP32 NFFT sizes per node set (device, np and the angular grid fix the
// nodes), measured once per process.
static std::mutex g_tune_mu;
static std::map<std::vector<double>, NfftTuning> g_tune;

static std::vector<double> tune_key(const nudim_params *p) {
  return {(double)p->device, (double)p->np, p->gam_min, p->gam_max, p->interval, p->padangle};
}

static bool autotune_enabled(const nudim_params *p) {
  const char *e = getenv("NUDIM_AUTOTUNE");
  return p->precision == NUDIM_FP32 && p->nfft_n <= 0 && p->nfft_w <= 0 && !(e && std::string(e) == "0");
}

// the tuned (n, W) of the job's node set, measuring it on first use
static NfftTuning tuned_nfft(const nudim_params *p, const Geometry &g, hipStream_t st) {
  const auto key = tune_key(p);
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find(key);
  if (it != g_tune.end()) return it->second;
  NfftTuning t = tune_nfft_fp32(p->np, g.x.data(), g.M, st, nullptr);
  g_tune[key] = t;
  return t;
}

// R from the known-node quadratic form (NUDIM_R_FORM=nfft: through the
// forward NFFT); see Job::rform
static bool rform_on() {
  const char *e = getenv("NUDIM_R_FORM");
  return !(e && std::string(e) == "nfft");
}

// Whether the projected-start solve keeps its ten search directions apart
// (bytes: their size) and the z transform applies f + sum_k alpha_k p_k,
// instead of f being read and written in nine row passes.  Up to 4 GiB of
// directions (3 GB at np = 384, nz = 256) this is 2-3 % faster per pass; at
// 24 GB (np = 768, nz = 512) the z transform sums its eleven streams at
// about 3.3 TB/s and the pass is only level with the per-step update, so
// larger jobs keep the per-step update and the memory.  NUDIM_KEEP_P=0 / 1
// turns it off / on whatever the size.
static bool keep_on(size_t bytes) {
  const char *e = getenv("NUDIM_KEEP_P");
  if (e && std::string(e) == "0") return false;
  if (e && std::string(e) == "1") return true;
  return bytes <= ((size_t)4 << 30);
}

static bool projected_start(int form, int N);

// Whether a job's passes run the NFFT.  With the projected CG start and R
// from the quadratic form they do not (only the load and the initial INFFT
// do): such a job always takes the default sizes, n = nice_fft_size(2 np)
// and 8 taps (what the measurement picks at sigma = 2 on the configured node
// sets), so its result does not depend on what the process measured before.
static bool passes_run_nfft(int form, int np) { return !(projected_start(form, np) && rform_on()); }

static void resolve_knobs(const nudim_params *p, int *n, int *W, int *form) {
  if (p->precision == NUDIM_FP64_PLAIN) {
    *n = p->nfft_n > 0 ? p->nfft_n : 2 * next_pow2(p->np);
    *W = p->nfft_w > 0 ? p->nfft_w : 14;
    *form = p->cg_form > 0 ? p->cg_form : NUDIM_CG_NFFT;
  } else {
    *n = p->nfft_n > 0 ? p->nfft_n : nice_fft_size(2 * p->np);
    *W = p->nfft_w > 0 ? p->nfft_w : 8;
    *form = p->cg_form > 0 ? p->cg_form : NUDIM_CG_TOEPLITZ;
  }
}

// The Toeplitz-form CGNR of a pass starts from z = A^H W (y - A f0) with
// y = K m + U A fh (the Fourier projector: measured values at the known
// nodes K, the current transform elsewhere).  Since A^H W K m is the same in
// every pass, z = b_K + T_U fh - T f0 with b_K = A^H W K m computed once,
// T_U = A^H W U A and T = A^H W A: two Toeplitz products on the fused passes
// instead of the adjoint NFFT and one product.  NUDIM_CG_START=nfft keeps the
// adjoint NFFT form.
static bool projected_start(int form, int N) {
  const char *e = getenv("NUDIM_CG_START");
  const bool off = e && std::string(e) == "nfft";
  return !off && form == NUDIM_CG_TOEPLITZ && g_use_fft2 && toep_supported(2 * N);
}

template <typename T>
struct Job : JobBase {
  nudim_params prm;
  const Geometry &geo;
  hipStream_t st;
  // P: planes (nz) of one volume; K volumes, Pt = K P Fourier planes; V and
  // Vt: voxels of one volume and of all; NN, MM: Fourier and sample sizes
  int N, P, M, n, W, form, K, Pt;
  long long V, Vt, NN, MM;
  bool adjoint_done = false;
  NfftOp<T> op;
  ToeplitzOp<T> toep;
  Cgnr<T> cg;
  DevBuf dw, dkind, dorder;
  DevBuf meas, y, Fh, F0;
  DevBuf img_a, img_b, img_c, grad, grad2;
  DevBuf d1, d2, tvpart_g, tvpart_a, tvsums;
  FftDesc fdz{};
  DevBuf twz;
  bool proj_start = false;  // projected_start(): b_K in bk
  DevBuf bk;
  // R from the quadratic form (with proj_start unless NUDIM_R_FORM=nfft):
  // ck = A_K^H m (unweighted), m2[p] = sum_K |m|^2, the column partials of
  // f^H T_K f from the CG start's two-input column pass, their sums and
  // Re <f, ck> per plane
  bool rform = false;
  DevBuf ck, m2, qpart, qsum, xr;
  DevBuf dknown;  // with proj_start: the known nodes (R needs A fh only there)
  int nknown = 0;
  static constexpr int kRBlocks = 1024;  // node strides of the fused R interpolation
  DevBuf rpart;                          // its block sums [2][kRBlocks][P]
  T *img = nullptr, *img_last = nullptr, *spare = nullptr;
  static constexpr int kTvBlocks = 1024;

  Job(const nudim_params &p, const Geometry &g, hipStream_t s, int nvol = 1) : prm(p), geo(g), st(s) {
    N = p.np;
    P = p.nz;
    K = nvol;
    Pt = K * P;
    M = g.M;
    resolve_knobs(&p, &n, &W, &form);
    if (autotune_enabled(&p) && passes_run_nfft(form, p.np)) {
      const NfftTuning t = tuned_nfft(&p, g, s);
      n = t.n;
      W = t.W;
    }
    V = (long long)N * N * P;
    Vt = V * K;
    NN = (long long)N * N * Pt;
    MM = (long long)M * Pt;
    op.init(N, n, W, g.x.data(), M, Pt);
    if (!g.w.empty()) {
      dw.alloc(g.w.size() * 8);
      NUDIM_HIP(hipMemcpy(dw.p, g.w.data(), g.w.size() * 8, hipMemcpyHostToDevice));
    }
    if (form == NUDIM_CG_TOEPLITZ) toep.init(N, g.x.data(), g.w.empty() ? nullptr : g.w.data(), M, Pt, st);
    cg.init(N, M, Pt, form, &op, form == NUDIM_CG_TOEPLITZ ? &toep : nullptr, dw.as<double>());
    proj_start = projected_start(form, N);
    // passes or solves that interpolate: the NFFT CGNR, and the pass without
    // the projected start (A fh for the projector)
    if (form != NUDIM_CG_TOEPLITZ || !proj_start) op.prepare_trafo();
    if (proj_start) {
      if (keep_on((size_t)10 * NN * sizeof(cx<T>))) cg.keep_directions(10);
      // kernel of T_U: the weights of the nodes the projector does not
      // replace by measured values (kind != 1)
      std::vector<double> wu(M);
      for (int m = 0; m < M; m++) wu[m] = g.kind[m] == 1 ? 0.0 : (g.w.empty() ? 1.0 : g.w[m]);
      toep.init_second(g.x.data(), wu.data(), M, st);
      bk.alloc(NN * sizeof(cx<T>));
      rpart.alloc((size_t)2 * kRBlocks * Pt * sizeof(double));
      std::vector<int> kn;
      for (int m = 0; m < M; m++)
        if (g.kind[m] == 1) kn.push_back(m);
      nknown = (int)kn.size();
      if (nknown > 0) {
        dknown.alloc(kn.size() * sizeof(int));
        NUDIM_HIP(hipMemcpy(dknown.p, kn.data(), kn.size() * sizeof(int), hipMemcpyHostToDevice));
      }
      rform = rform_on();
      if (rform) {
        // T_K: the kernel of unit weights on the known nodes
        std::vector<double> wk(M);
        for (int m = 0; m < M; m++) wk[m] = g.kind[m] == 1 ? 1.0 : 0.0;
        toep.init_rform(g.x.data(), wk.data(), M, st);
        ck.alloc(NN * sizeof(cx<T>));
        m2.alloc(Pt * sizeof(double));
        qpart.alloc((size_t)2 * N * Pt * sizeof(double));
        qsum.alloc(Pt * sizeof(double));
        xr.alloc(Pt * sizeof(double));
      }
    }
    dkind.alloc(M * sizeof(int));
    NUDIM_HIP(hipMemcpy(dkind.p, g.kind.data(), M * sizeof(int), hipMemcpyHostToDevice));
    dorder.alloc(M * sizeof(int));
    NUDIM_HIP(hipMemcpy(dorder.p, g.order.data(), M * sizeof(int), hipMemcpyHostToDevice));
    meas.alloc(MM * sizeof(cx<T>));
    y.alloc(MM * sizeof(cx<T>));
    Fh.alloc(NN * sizeof(cx<T>));
    F0.alloc(NN * sizeof(cx<T>));
    img_a.alloc(Vt * sizeof(T));
    img_b.alloc(Vt * sizeof(T));
    if (p.denoise && p.method == NUDIM_TV) {
      img_c.alloc(Vt * sizeof(T));
      grad.alloc(Vt * sizeof(T));
      grad2.alloc(Vt * sizeof(T));
      // zero outside the TV band for the life of the job
      NUDIM_HIP(hipMemset(grad.p, 0, Vt * sizeof(T)));
      NUDIM_HIP(hipMemset(grad2.p, 0, Vt * sizeof(T)));
    }
    d1.alloc(Pt * 8);
    d2.alloc(Pt * 8);
    tvpart_g.alloc((size_t)K * tv_blocks(N, N, P) * 8);
    tvpart_a.alloc((size_t)K * tv_blocks(N, N, P) * 8);
    tvsums.alloc(2 * K * 8);
    fdz = make_fft_desc(P);
    std::vector<double> t;
    fft_twiddles(P, &t);
    std::vector<cx<T>> h(P);
    for (int i = 0; i < P; i++) h[i] = mk((T)t[2 * i], (T)t[2 * i + 1]);
    twz.alloc(P * sizeof(cx<T>));
    NUDIM_HIP(hipMemcpy(twz.p, h.data(), P * sizeof(cx<T>), hipMemcpyHostToDevice));
    img = img_a.as<T>();
    img_last = img_b.as<T>();
    spare = img_c.as<T>();
    NUDIM_HIP(hipMemsetAsync(img, 0, Vt * sizeof(T), st));
  }

  // b_K = A^H W K m (the known nodes' measured values, weighted); with
  // rform also ck = A^H K m and m2 = sum_K |m|^2 per plane
  void known_adjoint() {
    NUDIM_CHECK(proj_start && bk.p, "b_K requested without the projected CG start");
    cx<T> *t = cg.t.template as<cx<T>>();
    hipLaunchKernelGGL(k_known_weighted<T>, dim3(cdiv(MM, 256)), dim3(256), 0, st, meas.as<cx<T>>(), dkind.as<int>(),
                       dw.as<double>(), t, MM, Pt);
    NUDIM_HIP(hipGetLastError());
    op.adjoint(t, bk.as<cx<T>>(), st);
    if (rform) {
      hipLaunchKernelGGL(k_known_weighted<T>, dim3(cdiv(MM, 256)), dim3(256), 0, st, meas.as<cx<T>>(),
                         dkind.as<int>(), (const double *)nullptr, t, MM, Pt);
      NUDIM_HIP(hipGetLastError());
      op.adjoint(t, ck.as<cx<T>>(), st);
      // |0 - m|^2 over the known nodes
      cx<T> *zero = cg.v.template as<cx<T>>();
      hipLaunchKernelGGL(k_zero_cx<T>, dim3(cdiv(MM, 256)), dim3(256), 0, st, zero, MM);
      cg.red.template rfactor<T>(zero, meas.as<cx<T>>(), dkind.as<int>(), M, m2.as<double>(), qsum.as<double>(), st);
    }
  }

  size_t bytes() const override {
    return op.bytes() + toep.bytes() + cg.bytes() + dw.bytes + dkind.bytes + dorder.bytes + meas.bytes + y.bytes +
           bk.bytes + dknown.bytes + rpart.bytes + ck.bytes + m2.bytes + qpart.bytes + qsum.bytes + xr.bytes +
           Fh.bytes + F0.bytes + img_a.bytes + img_b.bytes + img_c.bytes + grad.bytes + grad2.bytes + d1.bytes +
           d2.bytes + tvpart_g.bytes +
           tvpart_a.bytes + tvsums.bytes + twz.bytes;
  }

  void finish_load() override {
    if (proj_start && !adjoint_done) {
      known_adjoint();
      NUDIM_HIP(hipStreamSynchronize(st));
    }
    adjoint_done = true;
  }

  void load_meas_planes(const double *slices) override {
    NUDIM_CHECK(K == 1, "sample tables load into single-volume jobs only");
    DevBuf tmp((size_t)MM * 16);
    NUDIM_HIP(hipMemcpyAsync(tmp.p, slices, (size_t)MM * 16, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_cx_from_planes<T>, dim3(cdiv(MM, 256)), dim3(256), 0, st, tmp.as<double>(), meas.as<cx<T>>(),
                       (long long)M, P);
    NUDIM_HIP(hipGetLastError());
    adjoint_done = false;
    finish_load();
  }

  void load_projections(const double *proj, const double *rnd, int vol) override {
    NUDIM_CHECK(vol >= 0 && vol < K, "volume index out of range");
    int nx = prm.nx, nmeas = geo.nmeas;
    size_t pb = (size_t)nmeas * P * nx * 8;
    const long long Mv = (long long)M * P;  // samples of one volume
    DevBuf dproj(pb), dtomo((size_t)nmeas * P * N * sizeof(cx<double>)), drnd;
    NUDIM_HIP(hipMemcpyAsync(dproj.p, proj, pb, hipMemcpyHostToDevice, st));
    if (rnd) {
      drnd.alloc((size_t)Mv * 8);
      NUDIM_HIP(hipMemcpyAsync(drnd.p, rnd, (size_t)Mv * 8, hipMemcpyHostToDevice, st));
    }
    projection_ft_dev(nx, P, N, nmeas, dproj.as<double>(), dtomo.as<cx<double>>(), st);
    hipLaunchKernelGGL(k_assemble<T>, dim3(cdiv(Mv, 256)), dim3(256), 0, st, dtomo.as<cx<double>>(),
                       dkind.as<int>(), dorder.as<int>(), drnd.as<double>(), meas.as<cx<T>>(), M, N, P, Pt, vol * P);
    NUDIM_HIP(hipGetLastError());
    adjoint_done = false;
    if (K == 1) finish_load();
    NUDIM_HIP(hipStreamSynchronize(st));
  }

  // out = z IFFT of g (+ alpha p per plane when pv is non-null: the CG's last
  // f += alpha p, folded into the loads)
  // the z IFFT of g + sum_{s < nk} alpha_s p_s (p_s at pv + s N^2 P, alpha_s
  // per plane at alpha + s P)
  // (lines: N^2 K, line (y x) K + k of the Fourier side from volume k's
  // line (y x))
  void zfft_inverse(const cx<T> *g, T *out, const cx<T> *pv = nullptr, const double *alpha = nullptr, int nk = 1) {
    int Gs = line_chunk_log2(P, 1 << 30, (int)sizeof(cx<T>), FftEpt<T>::value);
    long long nlines = (long long)N * N * K;
    size_t lds = (size_t)P * ((1 << Gs) + 1) * sizeof(cx<T>);
    if (!pv) nk = 0;
    ProfScope ps(KC_ZFFT, st, (double)Vt * sizeof(cx<T>) * (1.0 + nk) + (double)Vt * sizeof(T));
    if (zfft2_launch<T>(P, true, nlines, nullptr, g, pv, alpha, nullptr, nullptr, out, twz.as<cx<T>>(), st, nk,
                        NN, K, V)) {
      NUDIM_HIP(hipGetLastError());
      return;
    }
    for (int s = 0; s < nk; s++)
      hipLaunchKernelGGL(k_axpy_planes<T>, dim3(cdiv(NN, 256)), dim3(256), 0, st, const_cast<cx<T> *>(g),
                         pv + (long long)s * NN, alpha + (long long)s * Pt, 1.0, NN, Pt);
    hipLaunchKernelGGL(k_zfft_inv<T>, dim3(cdiv(nlines, 1 << Gs)), dim3(kFftThreads), lds, st, g, out, fdz,
                       twz.as<cx<T>>(), Gs, nlines, K, V);
    NUDIM_HIP(hipGetLastError());
  }

  void zfft_forward(const T *in, cx<T> *fh, cx<T> *f0) {
    int Gs = line_chunk_log2(P, 1 << 30, (int)sizeof(cx<T>), FftEpt<T>::value);
    long long nlines = (long long)N * N * K;
    size_t lds = (size_t)P * ((1 << Gs) + 1) * sizeof(cx<T>);
    ProfScope ps(KC_ZFFT, st, (double)Vt * sizeof(T) + (f0 ? 2.0 : 1.0) * Vt * sizeof(cx<T>));
    if (zfft2_launch<T>(P, false, nlines, in, nullptr, nullptr, nullptr, fh, f0, nullptr, twz.as<cx<T>>(), st, 0, 0,
                        K, V)) {
      NUDIM_HIP(hipGetLastError());
      return;
    }
    hipLaunchKernelGGL(k_zfft_fwd<T>, dim3(cdiv(nlines, 1 << Gs)), dim3(kFftThreads), lds, st, in, fh, f0, fdz,
                       twz.as<cx<T>>(), Gs, nlines, K, V);
    NUDIM_HIP(hipGetLastError());
  }

  void initial() override {
    hipLaunchKernelGGL(k_zero_cx<T>, dim3(cdiv(NN, 256)), dim3(256), 0, st, F0.as<cx<T>>(), NN);
    cg.solve(meas.as<cx<T>>(), F0.as<cx<T>>(), 10, true, st);
    zfft_inverse(F0.as<cx<T>>(), img);
    NUDIM_HIP(hipStreamSynchronize(st));
  }

  void tv(T *u) {
    int ylo = 0, yhi = N - 1;
    if (prm.config_height && !prm.tv_full) {
      ylo = std::max(ylo, (2 * prm.ncenter - prm.nheight) / 2);
      yhi = std::min(yhi, (2 * prm.ncenter + prm.nheight) / 2);
    }
    if (yhi < ylo) return;
    // total_variation_discrete(_height)(img, 0.2f, 40, ...), reconstruction.c:630-632
    T *r = tv_sweeps<T>(u, spare, grad.as<T>(), grad2.as<T>(), N, N, P, ylo, yhi, 0.2f, 40, tvpart_g.as<double>(),
                        tvpart_a.as<double>(), tvsums.as<double>(), tv_blocks(N, N, P), st, K);
    if (r != img) std::swap(img, spare);
  }

  void diffusion(T *u) {
    // aniso_diffusion_2d(img, 0.25f, 0.005f, 80, ...), reconstruction.c:627
    // (reconstruction-parallel.c:764 passes 0.4f)
    const double lam = (double)(prm.diffusion_lambda > 0 ? prm.diffusion_lambda : 0.25f), kappa = (double)0.005f;
    diffusion_sweeps<T>(u, N, N, P, 80, lam, kappa * kappa, st, K);
  }

  void copy_to_host_ref(const T *u, double *out) {
    DevBuf tmp(V * 8);
    hipLaunchKernelGGL(k_vol_to_ref<T>, dim3(cdiv(V, 256)), dim3(256), 0, st, u, tmp.as<double>(), N, N, P);
    NUDIM_HIP(hipMemcpyAsync(out, tmp.p, V * 8, hipMemcpyDeviceToHost, st));
    NUDIM_HIP(hipStreamSynchronize(st));
  }

  void iterate(int nn, double *rslot) override {
    std::swap(img, img_last);  // img_last = img (reconstruction.c:551)
    zfft_forward(img_last, Fh.as<cx<T>>(), F0.as<cx<T>>());
    cx<T> *v = cg.v.template as<cx<T>>();
    if (proj_start && rform) {
      // R from the quadratic form: Re <fh, A_K^H m> here, fh^H T_K fh from
      // the CG start's column pass below
      cg.red.template redot<T>(Fh.as<cx<T>>(), ck.as<cx<T>>(), (long long)N * N, xr.as<double>(), st);
    } else if (proj_start) {
      // R needs A fh only at the known nodes: the interpolation sums the R
      // terms itself (no A fh stored)
      double *p1 = rpart.as<double>(), *p2 = p1 + (size_t)kRBlocks * Pt;
      op.trafo_rsums(Fh.as<cx<T>>(), meas.as<cx<T>>(), dknown.as<int>(), nknown, p1, p2, kRBlocks, st);
      hipLaunchKernelGGL(k_sum_partials, dim3(cdiv(Pt, 8)), dim3(256), 0, st, p1, kRBlocks, Pt, d1.as<double>());
      hipLaunchKernelGGL(k_sum_partials, dim3(cdiv(Pt, 8)), dim3(256), 0, st, p2, kRBlocks, Pt, d2.as<double>());
    } else {
      op.trafo(Fh.as<cx<T>>(), v, st);
      cg.red.template rfactor<T>(v, meas.as<cx<T>>(), dkind.as<int>(), M, d1.as<double>(), d2.as<double>(), st);
      hipLaunchKernelGGL(k_projector<T>, dim3(cdiv(MM, 256)), dim3(256), 0, st, v, meas.as<cx<T>>(),
                         dkind.as<int>(), y.as<cx<T>>(), MM, Pt);
    }
    if (!rform) {
      hipLaunchKernelGGL(k_rfactor_final, dim3(K), dim3(256), 0, st, d1.as<double>(), d2.as<double>(), P, rslot);
      NUDIM_HIP(hipGetLastError());
    }
    if (proj_start) {
      // z = b_K + T_U Fh - T f0, then the CG steps
      toep.apply_start(Fh.as<cx<T>>(), F0.as<cx<T>>(), bk.as<cx<T>>(), cg.z.template as<cx<T>>(),
                       cg.rowpart.template as<double>(), st, rform ? qpart.as<double>() : nullptr);
      if (rform) {
        hipLaunchKernelGGL(k_sum_partials, dim3(cdiv(Pt, 8)), dim3(256), 0, st, qpart.as<double>(), 2 * N, Pt,
                           qsum.as<double>());
        hipLaunchKernelGGL(k_rform_final, dim3(cdiv(Pt, 256)), dim3(256), 0, st, qsum.as<double>(), xr.as<double>(),
                           m2.as<double>(), d1.as<double>(), d2.as<double>(), Pt);
        hipLaunchKernelGGL(k_rfactor_final, dim3(K), dim3(256), 0, st, d1.as<double>(), d2.as<double>(), P, rslot);
        NUDIM_HIP(hipGetLastError());
      }
      // f += sum_k alpha_k p_k happens in the z transform's loads (f itself
      // is not needed afterwards: the next pass rebuilds f0); without kept
      // directions only the last step's update is left there
      cg.solve_from_residual(F0.as<cx<T>>(), 10, st, true, true);
      if (cg.kept >= 10)
        zfft_inverse(F0.as<cx<T>>(), img, cg.pk.template as<cx<T>>(), cg.alk.template as<double>(), 10);
      else
        zfft_inverse(F0.as<cx<T>>(), img, cg.pv.template as<cx<T>>(), cg.alpha.template as<double>());
    } else {
      cg.solve(y.as<cx<T>>(), F0.as<cx<T>>(), 10, false, st);
      zfft_inverse(F0.as<cx<T>>(), img);
    }
    if (prm.denoise) {
      if (nn == dump_nn && dump_before) copy_to_host_ref(img, dump_before);  // (dumps: nvol = 1)
      if (prm.method == NUDIM_DIFFUSION)
        diffusion(img);
      else
        tv(img);
      if (nn == dump_nn && dump_after) copy_to_host_ref(img, dump_after);
    }
    hipLaunchKernelGGL(k_constraint<T>, dim3(cdiv((size_t)N * P, 256), N, K), dim3(256), 0, st, img, img_last, N, P,
                       prm.nx, nn < prm.nhio ? 1 : 0, prm.config_height, prm.nheight, prm.ncenter);
    NUDIM_HIP(hipGetLastError());
  }

  void get_volume(double *out, int vol) override {
    NUDIM_CHECK(vol >= 0 && vol < K, "volume index out of range");
    copy_to_host_ref(img + (long long)vol * V, out);
  }

  void get_state(void *s[3]) const override {
    s[0] = img;
    s[1] = img_last;
    s[2] = spare;
  }
  void set_state(void *const s[3]) override {
    img = (T *)s[0];
    img_last = (T *)s[1];
    spare = (T *)s[2];
  }

  void set_volume(const double *in) override {
    NUDIM_CHECK(K == 1, "set_volume applies to single-volume jobs");
    DevBuf tmp(V * 8);
    NUDIM_HIP(hipMemcpyAsync(tmp.p, in, V * 8, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_vol_from_ref<T>, dim3(cdiv(V, 256)), dim3(256), 0, st, tmp.as<double>(), img, N, N, P);
    NUDIM_HIP(hipGetLastError());
    NUDIM_HIP(hipStreamSynchronize(st));
  }
};

// One pass captured as a hipGraph.  A pass's launches depend on the host
// state only through the three volume pointers and the HIO/ER switch, so a
// graph captured from one state replays any later pass that starts in it;
// the pointers cycle with period at most 6, so a job holds at most 12.
struct PassGraph {
  void *pre[3], *post[3];
  int hio;
  hipGraphExec_t exec;
};

struct nudim_ctx {
  nudim_params p{};
  Geometry geo;
  hipStream_t st = nullptr;
  std::unique_ptr<JobBase> job;
  int nvol = 1;              // volumes of the job (a batch when > 1)
  std::vector<char> vloaded;  // which volumes have projections
  int nn = 0;
  bool loaded = false, initialized = false;
  DevBuf rtrace;
  int rcap = 0;
  DevBuf rcur;  // R of a replayed pass (nvol values), copied to its trace slot after the graph
  std::vector<PassGraph> graphs;
  KernelProfiler prof;
  // also the clean-up of a context whose creation failed part way
  ~nudim_ctx() {
    (void)hipSetDevice(p.device);
    if (st) (void)hipStreamSynchronize(st);
    for (auto &g : graphs) (void)hipGraphExecDestroy(g.exec);
    graphs.clear();
    job.reset();
    if (st) (void)hipStreamDestroy(st);
  }
};

// pass nn through the graph captured for the job's current state
static void iterate_graph(nudim_ctx *c, int nn, double *rslot) {
  void *pre[3];
  c->job->get_state(pre);
  const int hio = nn < c->p.nhio ? 1 : 0;
  const PassGraph *hit = nullptr;
  for (const auto &g : c->graphs)
    if (g.hio == hio && std::equal(pre, pre + 3, g.pre)) hit = &g;
  if (!hit) {
    if (!c->rcur.p) c->rcur.alloc(c->nvol * sizeof(double));
    PassGraph g{};
    std::copy(pre, pre + 3, g.pre);
    g.hio = hio;
    hipGraph_t graph = nullptr;
    NUDIM_HIP(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
    try {
      c->job->iterate(nn, c->rcur.as<double>());
    } catch (...) {
      (void)hipStreamEndCapture(c->st, &graph);
      if (graph) (void)hipGraphDestroy(graph);
      c->job->set_state(pre);
      throw;
    }
    NUDIM_HIP(hipStreamEndCapture(c->st, &graph));
    const hipError_t e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
      c->job->set_state(pre);
      throw nudim::Error(NUDIM_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    c->job->get_state(g.post);
    c->graphs.push_back(g);
    hit = &c->graphs.back();
  }
  NUDIM_HIP(hipGraphLaunch(hit->exec, c->st));
  c->job->set_state(hit->post);
  NUDIM_HIP(hipMemcpyAsync(rslot, c->rcur.p, c->nvol * sizeof(double), hipMemcpyDeviceToDevice, c->st));
}

// makes the context's profiler current for the calling thread
struct ProfBind {
  KernelProfiler *old;
  explicit ProfBind(nudim_ctx *c) : old(g_prof) { g_prof = c->prof.cls ? &c->prof : nullptr; }
  ~ProfBind() { g_prof = old; }
};

extern "C" void nudim_default_params(nudim_params *p) {
  std::memset(p, 0, sizeof(*p));
  p->nx = p->nz = p->np = -1;
  p->gam_min = p->gam_max = 100;
  p->interval = -1;
  p->padangle = -1;
  p->nheight = p->ncenter = -1;
  p->nhio = p->ner = -1;
  p->method = NUDIM_TV;
  p->precision = NUDIM_FP64_PLAIN;
}

static int create_job(const nudim_params *p, int nvol, nudim_ctx **out) {
  API_BEGIN
  NUDIM_CHECK(out, "null output pointer");
  *out = nullptr;
  validate(p);
  NUDIM_CHECK(nvol >= 1 && nvol <= 4096, "a batch holds 1 to 4096 volumes");
  if (nvol > 1) {
    // the largest single jobs run (np = 768, nz = 512: n^2 nz = 1.2e9) stay
    // below 2^31 grid elements; a batch is held to the same bound
    int n, W, form;
    resolve_knobs(p, &n, &W, &form);
    NUDIM_CHECK((double)n * n * p->nz * nvol < 2147483648.0,
                "a batch's NFFT grid (n^2 nz nvol = " + std::to_string((double)n * n * p->nz * nvol) +
                    " elements) must stay below 2^31");
  }
  std::unique_ptr<nudim_ctx> c(new nudim_ctx());
  c->p = *p;
  c->nvol = nvol;
  c->vloaded.assign(nvol, 0);
  NUDIM_HIP(hipSetDevice(p->device));
  bool masked = false;
  for (unsigned w : p->cu_mask) masked |= w != 0;
  // a CU-masked stream cannot be created non-blocking; the library's
  // synchronous copies use the per-thread default stream, so blocking
  // streams do not serialise with them
  if (masked)
    NUDIM_HIP(hipExtStreamCreateWithCUMask(&c->st, 8, p->cu_mask));
  else
    NUDIM_HIP(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
  make_geometry(p, &c->geo);
  if (p->precision == NUDIM_FP64_PLAIN)
    c->job.reset(new Job<double>(c->p, c->geo, c->st, nvol));
  else
    c->job.reset(new Job<float>(c->p, c->geo, c->st, nvol));
  NUDIM_HIP(hipStreamSynchronize(c->st));
  *out = c.release();
  return NUDIM_OK;
  API_END
}

extern "C" int nudim_create(const nudim_params *p, nudim_ctx **out) { return create_job(p, 1, out); }

extern "C" int nudim_batch_create(const nudim_params *p, int nvol, nudim_ctx **out) {
  return create_job(p, nvol, out);
}

extern "C" void nudim_destroy(nudim_ctx *c) { delete c; }

extern "C" int nudim_sizes(const nudim_ctx *c, int *nmeas, int *M, int *nang) {
  if (!c) return fail(NUDIM_EINVAL, "null context");
  if (nmeas) *nmeas = c->geo.nmeas;
  if (M) *M = c->geo.M;
  if (nang) *nang = c->geo.nang;
  return NUDIM_OK;
}

static int load_volume(nudim_ctx *c, int vol, const double *proj) {
  API_BEGIN
  NUDIM_CHECK(c && proj, "null argument");
  NUDIM_CHECK(vol >= 0 && vol < c->nvol, "volume index out of range");
  NUDIM_HIP(hipSetDevice(c->p.device));
  std::vector<double> rnd;
  if (c->p.random) {
    // gsl_rng_uniform(rng)*10 for every unknown and held-out sample, planes
    // outer, nodes inner (reconstruction.c:437-469)
    Mt19937 mt(c->p.seed);
    rnd.assign((size_t)c->geo.M * c->p.nz, 0.0);
    for (int j = 0; j < c->p.nz; j++)
      for (int i = 0; i < c->geo.M; i++)
        if (c->geo.kind[i] != 1) rnd[(size_t)j * c->geo.M + i] = mt.uniform() * 10;
  }
  c->job->load_projections(proj, rnd.empty() ? nullptr : rnd.data(), vol);
  c->vloaded[vol] = 1;
  c->loaded = std::all_of(c->vloaded.begin(), c->vloaded.end(), [](char v) { return v != 0; });
  return NUDIM_OK;
  API_END
}

extern "C" int nudim_load_projections(nudim_ctx *c, const double *proj) {
  if (c && c->nvol > 1) return fail(NUDIM_ESTATE, "a batch loads each volume with nudim_batch_load_projections");
  return load_volume(c, 0, proj);
}

extern "C" int nudim_batch_load_projections(nudim_ctx *c, int vol, const double *proj) {
  return load_volume(c, vol, proj);
}

extern "C" int nudim_load_samples(nudim_ctx *c, const double *slices, const int *kind) {
  API_BEGIN
  NUDIM_CHECK(c && slices, "null argument");
  NUDIM_HIP(hipSetDevice(c->p.device));
  if (kind) {
    for (int i = 0; i < c->geo.M; i++)
      NUDIM_CHECK(kind[i] == c->geo.kind[i], "sample classes differ from the job geometry");
  }
  c->job->load_meas_planes(slices);
  c->loaded = true;
  return NUDIM_OK;
  API_END
}

extern "C" int nudim_initial(nudim_ctx *c) {
  API_BEGIN
  NUDIM_CHECK(c, "null context");
  if (!c->loaded) return fail(NUDIM_ESTATE, "load projections (of every volume) before nudim_initial");
  NUDIM_HIP(hipSetDevice(c->p.device));
  c->job->finish_load();
  c->job->initial();
  c->initialized = true;
  c->nn = 0;
  return NUDIM_OK;
  API_END
}

extern "C" int nudim_iterate(nudim_ctx *c, int count, double *rfactor) {
  API_BEGIN
  NUDIM_CHECK(c && count >= 0, "invalid argument");
  if (!c->initialized) return fail(NUDIM_ESTATE, "call nudim_initial (or nudim_set_volume) first");
  NUDIM_HIP(hipSetDevice(c->p.device));
  const int nr = count * c->nvol;  // R values: [count][nvol]
  if (c->rcap < nr) {
    c->rtrace.alloc((size_t)nr * 8);
    c->rcap = nr;
  }
  if (c->loaded) c->job->finish_load();  // volumes reloaded since the last pass
  {
    ProfBind pb(c);
    // passes that time kernels or dump volumes run their launches directly
    for (int i = 0; i < count; i++) {
      const int nn = c->nn + i;
      double *slot = c->rtrace.as<double>() + (size_t)i * c->nvol;
      if (c->p.use_graph && !c->prof.cls && nn != c->job->dump_nn)
        iterate_graph(c, nn, slot);
      else
        c->job->iterate(nn, slot);
    }
  }
  if (rfactor && count > 0)
    NUDIM_HIP(hipMemcpyAsync(rfactor, c->rtrace.p, (size_t)nr * 8, hipMemcpyDeviceToHost, c->st));
  NUDIM_HIP(hipStreamSynchronize(c->st));
  c->prof.collect();
  c->nn += count;
  return NUDIM_OK;
  API_END
}

extern "C" int nudim_get_volume(nudim_ctx *c, double *img) {
  if (c && c->nvol > 1) return fail(NUDIM_ESTATE, "a batch returns each volume with nudim_batch_get_volume");
  return nudim_batch_get_volume(c, 0, img);
}

extern "C" int nudim_batch_get_volume(nudim_ctx *c, int vol, double *img) {
  API_BEGIN
  NUDIM_CHECK(c && img, "null argument");
  NUDIM_CHECK(vol >= 0 && vol < c->nvol, "volume index out of range");
  NUDIM_HIP(hipSetDevice(c->p.device));
  c->job->get_volume(img, vol);
  return NUDIM_OK;
  API_END
}

extern "C" int nudim_set_volume(nudim_ctx *c, const double *img, int nn) {
  API_BEGIN
  NUDIM_CHECK(c && img && nn >= 0, "invalid argument");
  if (c->nvol > 1) return fail(NUDIM_ESTATE, "nudim_set_volume applies to single-volume jobs");
  NUDIM_HIP(hipSetDevice(c->p.device));
  c->job->set_volume(img);
  c->nn = nn;
  c->initialized = true;
  return NUDIM_OK;
  API_END
}

extern "C" int nudim_set_dump(nudim_ctx *c, int nn, double *before, double *after) {
  if (!c) return fail(NUDIM_EINVAL, "null context");
  if (c->nvol > 1) return fail(NUDIM_ESTATE, "volume dumps apply to single-volume jobs");
  c->job->dump_nn = nn;
  c->job->dump_before = before;
  c->job->dump_after = after;
  return NUDIM_OK;
}

extern "C" int nudim_profile(nudim_ctx *c, int kernel_class) {
  if (!c || kernel_class < 0 || kernel_class > KC_DIFFUSION) return fail(NUDIM_EINVAL, "invalid profile request");
  c->prof.reset(kernel_class);
  return NUDIM_OK;
}

extern "C" int nudim_profile_read(nudim_ctx *c, double *ms, long *launches, double *bytes) {
  if (!c) return fail(NUDIM_EINVAL, "null context");
  if (ms) *ms = c->prof.ms;
  if (launches) *launches = c->prof.launches;
  if (bytes) *bytes = c->prof.bytes;
  return NUDIM_OK;
}

extern "C" size_t nudim_device_bytes(const nudim_ctx *c) { return c && c->job ? c->job->bytes() : 0; }

extern "C" int nudim_synchronize(nudim_ctx *c) {
  API_BEGIN
  NUDIM_CHECK(c, "null context");
  NUDIM_HIP(hipStreamSynchronize(c->st));
  return NUDIM_OK;
  API_END
}

extern "C" int nudim_estimate_bytes(const nudim_params *p, size_t *bytes) {
  API_BEGIN
  validate(p);
  NUDIM_CHECK(bytes, "null output");
  Geometry g;
  make_geometry(p, &g);
  int n, W, form;
  resolve_knobs(p, &n, &W, &form);
  size_t cb = p->precision == NUDIM_FP64_PLAIN ? 16 : 8, rb = cb / 2;
  size_t N = p->np, P = p->nz, M = g.M;
  size_t NN = N * N * P, MM = M * P;
  size_t b = 0;
  b += (size_t)N * n * P * cb + (size_t)n * n * P * cb;           // NFFT X, grid
  b += (size_t)M * W * W * (4 + rb) + (size_t)n * n * 4 + M * W * 16;  // gather maps, window tables
  {
    // patch lists: at most one entry per pair of patch bands a window meets
    auto entries = [&](int pr, int pc) { return M * (size_t)((W + pr - 2) / pr + 1) * ((W + pc - 2) / pc + 1); };
    const int gr = kGatherR, gc = kGatherC, pr = kPatchR, pc = kPatchC;
    const size_t grec = cb == 16 ? sizeof(PatchRec<double, kGatherR, kGatherC>) : sizeof(PatchRec<float, kGatherR, kGatherC>);
    const size_t rec = cb == 16 ? sizeof(PatchRec<double>) : sizeof(PatchRec<float>);
    // the patch tables belong to the geometry (built at every plane count
    // where the grid allows them); the gather's segment sums (split patches:
    // at most 2 E / 128 of gr gc cells) and the interpolation's partial sums
    // (only for jobs that interpolate: Job's prepare_trafo) to plane counts
    // that are multiples of 64
    const bool interp = form != NUDIM_CG_TOEPLITZ || !projected_start(form, (int)N);
    if (n % gr == 0 && n % gc == 0 && n >= W + std::max(gr, gc)) {
      // records, node lists, segments (padded to the XCD sectors: at most twice the patches)
      b += entries(gr, gc) * (grec + 16) + (size_t)2 * n * n / (gr * gc) * 16 + (size_t)M * 4;
      if (P % 64 == 0) b += 2 * entries(gr, gc) / 128 * gr * gc * P * cb;
      if (n % (2 * pc) == 0 && n % (2 * pr) == 0 && n >= W + 2 * pc) {
        b += entries(2 * pr, 2 * pc) * (8 + 4 * rec) + (size_t)n * n / (4 * pr * pc) * 16;
        if (P % 64 == 0 && interp) b += entries(2 * pr, 2 * pc) * P * cb;
      }
    }
  }
  if (form == NUDIM_CG_TOEPLITZ) b += (size_t)4 * N * N * (cb + rb) + (size_t)N * 2 * N * P * cb + NN * cb + 2 * N * P * 8;
  if (projected_start(form, (int)N))  // T_U, X2, b_K, R partials, known list; c_K and the quadratic form's partials
    b += (size_t)4 * N * N * rb + (size_t)N * 2 * N * P * cb + NN * cb + (size_t)2 * 1024 * P * 8 + M * 4 +
         NN * cb + (size_t)2 * N * P * 8;
  if (projected_start(form, (int)N) && keep_on(10 * NN * cb)) b += 10 * NN * cb + (size_t)10 * P * 8;  // kept directions
  b += 2 * NN * cb + 3 * MM * cb;                                   // CG z, p; r, v, t
  b += 2 * MM * cb + 2 * NN * cb;                                   // meas, y, Fh, F0
  b += 2 * NN * rb + ((p->denoise && p->method == NUDIM_TV) ? 3 * NN * rb : 0);  // volumes, TV gradients
  *bytes = b;
  return NUDIM_OK;
  API_END
}

// ---------------------------------------------------------------------------
// stateless parity entry points
// ---------------------------------------------------------------------------

struct TmpStream {
  hipStream_t s = nullptr;
  TmpStream() { NUDIM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)); }
  ~TmpStream() {
    if (s) {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
  }
};

extern "C" int nudim_autotune(const nudim_params *p, int *n, int *W) {
  API_BEGIN
  validate(p);
  NUDIM_CHECK(n && W, "null output");
  int form;
  resolve_knobs(p, n, W, &form);
  if (autotune_enabled(p) && passes_run_nfft(form, p->np)) {
    NUDIM_HIP(hipSetDevice(p->device));
    Geometry g;
    make_geometry(p, &g);
    TmpStream ts;
    const NfftTuning t = tuned_nfft(p, g, ts.s);
    *n = t.n;
    *W = t.W;
  }
  return NUDIM_OK;
  API_END
}


template <typename T>
static void to_dev_planes(const double *h, long long K, int P, DevBuf &out, hipStream_t st) {
  DevBuf tmp((size_t)K * P * 16);
  NUDIM_HIP(hipMemcpyAsync(tmp.p, h, (size_t)K * P * 16, hipMemcpyHostToDevice, st));
  out.alloc((size_t)K * P * sizeof(cx<T>));
  hipLaunchKernelGGL(k_cx_from_planes<T>, dim3(cdiv(K * P, 256)), dim3(256), 0, st, tmp.as<double>(), out.as<cx<T>>(),
                     K, P);
  NUDIM_HIP(hipGetLastError());
  NUDIM_HIP(hipStreamSynchronize(st));
}

template <typename T>
static void from_dev_planes(const DevBuf &d, long long K, int P, double *h, hipStream_t st) {
  DevBuf tmp((size_t)K * P * 16);
  hipLaunchKernelGGL(k_cx_to_planes<T>, dim3(cdiv(K * P, 256)), dim3(256), 0, st, d.as<cx<T>>(), tmp.as<double>(), K,
                     P);
  NUDIM_HIP(hipGetLastError());
  NUDIM_HIP(hipMemcpyAsync(h, tmp.p, (size_t)K * P * 16, hipMemcpyDeviceToHost, st));
  NUDIM_HIP(hipStreamSynchronize(st));
}

template <typename T>
static void nfft_parity(int N, int M, const double *x, int n, int W, int P, const double *in, double *out, bool adj) {
  TmpStream ts;
  NfftOp<T> op;
  op.init(N, n, W, x, M, P);
  DevBuf a, b((size_t)(adj ? (long long)N * N : M) * P * sizeof(cx<T>));
  if (!adj) {
    to_dev_planes<T>(in, (long long)N * N, P, a, ts.s);
    op.trafo(a.as<cx<T>>(), b.as<cx<T>>(), ts.s);
    from_dev_planes<T>(b, M, P, out, ts.s);
  } else {
    to_dev_planes<T>(in, M, P, a, ts.s);
    op.adjoint(a.as<cx<T>>(), b.as<cx<T>>(), ts.s);
    from_dev_planes<T>(b, (long long)N * N, P, out, ts.s);
  }
}

extern "C" int nudim_nfft_trafo(int precision, int N, int M, const double *x, int n, int W, int P,
                                const double *fhat, double *f) {
  API_BEGIN
  NUDIM_CHECK(x && fhat && f && N > 0 && M > 0 && P > 0, "invalid argument");
  if (precision == NUDIM_FP32) nfft_parity<float>(N, M, x, n, W, P, fhat, f, false);
  else nfft_parity<double>(N, M, x, n, W, P, fhat, f, false);
  return NUDIM_OK;
  API_END
}

extern "C" int nudim_nfft_adjoint(int precision, int N, int M, const double *x, int n, int W, int P,
                                  const double *f, double *fhat) {
  API_BEGIN
  NUDIM_CHECK(x && fhat && f && N > 0 && M > 0 && P > 0, "invalid argument");
  if (precision == NUDIM_FP32) nfft_parity<float>(N, M, x, n, W, P, f, fhat, true);
  else nfft_parity<double>(N, M, x, n, W, P, f, fhat, true);
  return NUDIM_OK;
  API_END
}

template <typename T>
static void cgnr_parity(int form, int N, int M, const double *x, const double *w, int n, int W, int P,
                        const double *y, double *fhat, int nsteps) {
  TmpStream ts;
  NfftOp<T> op;
  op.init(N, n, W, x, M, P);
  ToeplitzOp<T> toep;
  if (form == NUDIM_CG_TOEPLITZ) toep.init(N, x, w, M, P, ts.s);
  DevBuf dw;
  if (w) {
    dw.alloc((size_t)M * 8);
    NUDIM_HIP(hipMemcpy(dw.p, w, (size_t)M * 8, hipMemcpyHostToDevice));
  }
  Cgnr<T> cg;
  cg.init(N, M, P, form, &op, form == NUDIM_CG_TOEPLITZ ? &toep : nullptr, dw.as<double>());
  DevBuf dy, df;
  to_dev_planes<T>(y, M, P, dy, ts.s);
  to_dev_planes<T>(fhat, (long long)N * N, P, df, ts.s);
  cg.solve(dy.as<cx<T>>(), df.as<cx<T>>(), nsteps, false, ts.s);
  from_dev_planes<T>(df, (long long)N * N, P, fhat, ts.s);
}

extern "C" int nudim_cgnr(int precision, int form, int N, int M, const double *x, const double *w, int n, int W,
                          int P, const double *y, double *fhat, int nsteps) {
  API_BEGIN
  NUDIM_CHECK(x && y && fhat && N > 0 && M > 0 && P > 0 && nsteps >= 0, "invalid argument");
  if (form <= 0) form = precision == NUDIM_FP32 ? NUDIM_CG_TOEPLITZ : NUDIM_CG_NFFT;
  if (precision == NUDIM_FP32) cgnr_parity<float>(form, N, M, x, w, n, W, P, y, fhat, nsteps);
  else cgnr_parity<double>(form, N, M, x, w, n, W, P, y, fhat, nsteps);
  return NUDIM_OK;
  API_END
}

template <typename T>
static void zfft_parity(int len, int howmany, double *data, int sign) {
  TmpStream ts;
  LineFft<T> f;
  f.init(len);
  DevBuf a, tmp((size_t)len * howmany * 16);
  NUDIM_HIP(hipMemcpyAsync(tmp.p, data, (size_t)len * howmany * 16, hipMemcpyHostToDevice, ts.s));
  a.alloc((size_t)len * howmany * sizeof(cx<T>));
  // reference layout [len][howmany] is already "plane innermost" with P = howmany
  hipLaunchKernelGGL(k_cx_from_planes<T>, dim3(cdiv((long long)len * howmany, 256)), dim3(256), 0, ts.s,
                     tmp.as<double>(), a.as<cx<T>>(), (long long)len * howmany, 1);
  LineIO io{0, howmany, len, 0, nullptr, nullptr};
  f.run(a.as<cx<T>>(), a.as<cx<T>>(), 1, io, io, sign < 0 ? -1 : +1, howmany, ts.s);
  hipLaunchKernelGGL(k_cx_to_planes<T>, dim3(cdiv((long long)len * howmany, 256)), dim3(256), 0, ts.s,
                     a.as<cx<T>>(), tmp.as<double>(), (long long)len * howmany, 1);
  NUDIM_HIP(hipGetLastError());
  NUDIM_HIP(hipMemcpyAsync(data, tmp.p, (size_t)len * howmany * 16, hipMemcpyDeviceToHost, ts.s));
  NUDIM_HIP(hipStreamSynchronize(ts.s));
}

extern "C" int nudim_zfft(int precision, int len, int howmany, double *data, int sign) {
  API_BEGIN
  NUDIM_CHECK(data && len > 0 && howmany > 0, "invalid argument");
  if (precision == NUDIM_FP32) zfft_parity<float>(len, howmany, data, sign);
  else zfft_parity<double>(len, howmany, data, sign);
  return NUDIM_OK;
  API_END
}

// Volume-level parity: build a throw-away job-shaped state without geometry.
template <typename T>
struct VolTmp {
  TmpStream ts;
  DevBuf u, part_g, part_a, sums;
  int nz, ny, nx;
  long long V;
  VolTmp(const double *h, int nz_, int ny_, int nx_) : nz(nz_), ny(ny_), nx(nx_) {
    V = (long long)nz * ny * nx;
    DevBuf tmp(V * 8);
    NUDIM_HIP(hipMemcpyAsync(tmp.p, h, V * 8, hipMemcpyHostToDevice, ts.s));
    u.alloc(V * sizeof(T));
    hipLaunchKernelGGL(k_vol_from_ref<T>, dim3(cdiv(V, 256)), dim3(256), 0, ts.s, tmp.as<double>(), u.as<T>(), ny, nx,
                       nz);
    NUDIM_HIP(hipGetLastError());
    NUDIM_HIP(hipStreamSynchronize(ts.s));
  }
  void out(double *h) {
    DevBuf tmp(V * 8);
    hipLaunchKernelGGL(k_vol_to_ref<T>, dim3(cdiv(V, 256)), dim3(256), 0, ts.s, u.as<T>(), tmp.as<double>(), ny, nx,
                       nz);
    NUDIM_HIP(hipGetLastError());
    NUDIM_HIP(hipMemcpyAsync(h, tmp.p, V * 8, hipMemcpyDeviceToHost, ts.s));
    NUDIM_HIP(hipStreamSynchronize(ts.s));
  }
  void tv(float lam, int niter, int ylo, int yhi) {
    ylo = std::max(ylo, 0);
    yhi = std::min(yhi, ny - 1);
    if (yhi < ylo) return;
    DevBuf u2(V * sizeof(T));
    const int nb = tv_blocks(ny, nx, nz);
    part_g.alloc(nb * 8);
    part_a.alloc(nb * 8);
    sums.alloc(16);
    DevBuf g0(V * sizeof(T)), g1(V * sizeof(T));
    NUDIM_HIP(hipMemsetAsync(g0.p, 0, V * sizeof(T), ts.s));
    NUDIM_HIP(hipMemsetAsync(g1.p, 0, V * sizeof(T), ts.s));
    T *r = tv_sweeps<T>(u.as<T>(), u2.as<T>(), g0.as<T>(), g1.as<T>(), ny, nx, nz, ylo, yhi, lam, niter,
                        part_g.as<double>(), part_a.as<double>(), sums.as<double>(), nb, ts.s);
    if (r != u.as<T>()) std::swap(u, u2);
    NUDIM_HIP(hipStreamSynchronize(ts.s));
  }
};

extern "C" int nudim_tv(int precision, double *u, int nz, int ny, int nx, float lam, int niter) {
  API_BEGIN
  NUDIM_CHECK(u && nz > 0 && ny > 0 && nx > 0 && niter >= 0, "invalid argument");
  if (precision == NUDIM_FP32) {
    VolTmp<float> v(u, nz, ny, nx);
    v.tv(lam, niter, 0, ny - 1);
    v.out(u);
  } else {
    VolTmp<double> v(u, nz, ny, nx);
    v.tv(lam, niter, 0, ny - 1);
    v.out(u);
  }
  return NUDIM_OK;
  API_END
}

extern "C" int nudim_tv_height(int precision, double *u, int nz, int ny, int nx, float lam, int niter, int nheight,
                               int ncenter) {
  API_BEGIN
  NUDIM_CHECK(u && nz > 0 && ny > 0 && nx > 0 && niter >= 0, "invalid argument");
  int ylo = (2 * ncenter - nheight) / 2, yhi = (2 * ncenter + nheight) / 2;
  if (precision == NUDIM_FP32) {
    VolTmp<float> v(u, nz, ny, nx);
    v.tv(lam, niter, ylo, yhi);
    v.out(u);
  } else {
    VolTmp<double> v(u, nz, ny, nx);
    v.tv(lam, niter, ylo, yhi);
    v.out(u);
  }
  return NUDIM_OK;
  API_END
}

// ---------------------------------------------------------------------------
// bandwidth probes: plain 16-byte vector loads and stores over a buffer of
// known size (the reference point of achieved GB/s and of the HBM traffic
// counters)
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(256) k_stream_read(const float4 *__restrict__ a, long long n4, float *__restrict__ out) {
  float s = 0.f;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (long long)gridDim.x * blockDim.x) {
    const float4 v = a[i];
    s += v.x + v.y + v.z + v.w;
  }
  __shared__ float red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) k_stream_copy(const float4 *__restrict__ a, float4 *__restrict__ b, long long n4) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (long long)gridDim.x * blockDim.x)
    b[i] = a[i];
}

extern "C" int nudim_stream_probe(int device, size_t bytes, int kind, int reps, double *gbs) {
  API_BEGIN
  NUDIM_CHECK(gbs && bytes >= 16 && reps >= 1 && (kind == 0 || kind == 1), "invalid argument");
  NUDIM_HIP(hipSetDevice(device));
  const long long n4 = (long long)(bytes / 16);
  DevBuf a((size_t)n4 * 16), b(kind == 1 ? (size_t)n4 * 16 : 0), out(2048 * 4 * 8);
  TmpStream ts;
  NUDIM_HIP(hipMemsetAsync(a.p, 0, (size_t)n4 * 16, ts.s));
  hipEvent_t e0, e1;
  NUDIM_HIP(hipEventCreate(&e0));
  NUDIM_HIP(hipEventCreate(&e1));
  const unsigned blocks = 256 * 8 * 4;  // 32 blocks of 256 threads per CU
  double best = 0;
  for (int r = 0; r < reps; r++) {
    NUDIM_HIP(hipEventRecord(e0, ts.s));
    if (kind == 0)
      hipLaunchKernelGGL(k_stream_read, dim3(blocks), dim3(256), 0, ts.s, a.as<float4>(), n4, out.as<float>());
    else
      hipLaunchKernelGGL(k_stream_copy, dim3(blocks), dim3(256), 0, ts.s, a.as<float4>(), b.as<float4>(), n4);
    NUDIM_HIP(hipGetLastError());
    NUDIM_HIP(hipEventRecord(e1, ts.s));
    NUDIM_HIP(hipEventSynchronize(e1));
    float ms = 0;
    NUDIM_HIP(hipEventElapsedTime(&ms, e0, e1));
    const double moved = (kind == 0 ? 1.0 : 2.0) * (double)n4 * 16;
    best = std::max(best, moved / (ms * 1e-3) / 1e9);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *gbs = best;
  return NUDIM_OK;
  API_END
}

// ---------------------------------------------------------------------------
// TV-height on caller device memory with a file-scope scratch array (no
// device allocation): for callers that cannot call hipMalloc
// ---------------------------------------------------------------------------

// Scratch of nudim_tv_height_noalloc: the device-layout volume u [y][x][z],
// the second u buffer, the two gradient buffers, the two partial-norm arrays
// and the norm pair.  Each GPU holds its own copy until the process ends.
static constexpr size_t kScratchBytes = (size_t)68 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// its device address on each GPU (hipGetSymbolAddress, once per device)
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[64];

static unsigned char *scratch_on_current_device() {
  int dev = -1;
  NUDIM_HIP(hipGetDevice(&dev));
  NUDIM_CHECK(dev >= 0 && dev < 64, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    NUDIM_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    NUDIM_CHECK(ptr != nullptr, "hipGetSymbolAddress(scratch) returned a null address");
    g_scratch_ptr[dev] = ptr;
  }
  return static_cast<unsigned char *>(g_scratch_ptr[dev]);
}

static size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

__global__ void k_zero_words(uint4 *__restrict__ d, long long nwords) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nwords) d[i] = make_uint4(0u, 0u, 0u, 0u);
}

// scratch bytes the TV sweeps of one [nz][ny][nx] volume take: norm pair,
// two partial-sum arrays, four volume buffers
template <typename T>
static size_t tv_scratch_bytes(int nz, int ny, int nx) {
  const size_t V = (size_t)nz * ny * nx;
  return 256 + 2 * round256((size_t)tv_blocks(ny, nx, nz) * 8) + 4 * round256(V * sizeof(T));
}

template <typename T>
static void tv_height_noalloc(double *d_u, int nz, int ny, int nx, float lam, int niter, int ylo, int yhi,
                              hipStream_t st) {
  const size_t need = tv_scratch_bytes<T>(nz, ny, nx);
  if (need > kScratchBytes)
    throw Error(NUDIM_EINVAL, "nudim_tv_height_noalloc: a " + std::to_string(nz) + "x" + std::to_string(ny) + "x" +
                                  std::to_string(nx) + " volume needs " + std::to_string(need) +
                                  " bytes of scratch; the maximum is " + std::to_string(kScratchBytes));
  unsigned char *base = scratch_on_current_device();
  const long long V = (long long)nz * ny * nx;
  const int nb = tv_blocks(ny, nx, nz);
  const size_t vb = round256((size_t)V * sizeof(T)), pb = round256((size_t)nb * 8);
  double *sums = reinterpret_cast<double *>(base);
  double *part_g = reinterpret_cast<double *>(base + 256);
  double *part_a = reinterpret_cast<double *>(base + 256 + pb);
  unsigned char *vols = base + 256 + 2 * pb;
  T *u0 = reinterpret_cast<T *>(vols), *u1 = reinterpret_cast<T *>(vols + vb);
  T *g0 = reinterpret_cast<T *>(vols + 2 * vb), *g1 = reinterpret_cast<T *>(vols + 3 * vb);
  // the gradient buffers keep the previous call's values: clear both (they
  // must be zero outside the band rows)
  hipLaunchKernelGGL(k_zero_words, dim3(cdiv(2 * vb / 16, 256)), dim3(256), 0, st, reinterpret_cast<uint4 *>(g0),
                     (long long)(2 * vb / 16));
  hipLaunchKernelGGL(k_vol_from_ref<T>, dim3(cdiv(V, 256)), dim3(256), 0, st, (const double *)d_u, u0, ny, nx, nz);
  NUDIM_HIP(hipGetLastError());
  T *r = tv_sweeps<T>(u0, u1, g0, g1, ny, nx, nz, ylo, yhi, lam, niter, part_g, part_a, sums, nb, st);
  hipLaunchKernelGGL(k_vol_to_ref<T>, dim3(cdiv(V, 256)), dim3(256), 0, st, (const T *)r, d_u, ny, nx, nz);
  NUDIM_HIP(hipGetLastError());
}

extern "C" int nudim_tv_height_noalloc(int precision, double *d_u, int nz, int ny, int nx, float lam, int niter,
                                       int nheight, int ncenter, void *stream) {
  API_BEGIN
  NUDIM_CHECK(d_u && nz > 0 && ny > 0 && nx > 0 && niter >= 0, "invalid argument");
  NUDIM_CHECK(precision == NUDIM_FP32 || precision == NUDIM_FP64_PLAIN, "invalid precision");
  int ylo = (2 * ncenter - nheight) / 2, yhi = (2 * ncenter + nheight) / 2;
  ylo = std::max(ylo, 0);
  yhi = std::min(yhi, ny - 1);
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (precision == NUDIM_FP32) tv_height_noalloc<float>(d_u, nz, ny, nx, lam, niter, ylo, yhi, st);
  else tv_height_noalloc<double>(d_u, nz, ny, nx, lam, niter, ylo, yhi, st);
  return NUDIM_OK;
  API_END
}

template <typename T>
static void diffusion_parity(double *u, int nz, int ny, int nx, float lambda, float kappa, int niter) {
  VolTmp<T> v(u, nz, ny, nx);
  const double lam = (double)lambda, k = (double)kappa;
  diffusion_sweeps<T>(v.u.template as<T>(), ny, nx, nz, niter, lam, k * k, v.ts.s);
  v.out(u);
}

extern "C" int nudim_diffusion2d(int precision, double *u, int nz, int ny, int nx, float lambda, float kappa,
                                 int niter) {
  API_BEGIN
  NUDIM_CHECK(u && nz > 0 && ny > 0 && nx > 0 && niter >= 0, "invalid argument");
  NUDIM_CHECK(nz == nx, "aniso_diffusion_2d needs nz == nx (denoise.c:349-350 clamps x with the z extent)");
  if (precision == NUDIM_FP32) diffusion_parity<float>(u, nz, ny, nx, lambda, kappa, niter);
  else diffusion_parity<double>(u, nz, ny, nx, lambda, kappa, niter);
  return NUDIM_OK;
  API_END
}

template <typename T>
static void constraint_parity(double *img, const double *last, int nz, int np, int nx, int is_hio, int ch, int nh,
                              int nc) {
  VolTmp<T> a(img, nz, np, np), b(last, nz, np, np);
  NUDIM_HIP(hipStreamSynchronize(b.ts.s));
  hipLaunchKernelGGL(k_constraint<T>, dim3(cdiv((size_t)np * nz, 256), np, 1), dim3(256), 0, a.ts.s,
                     a.u.template as<T>(), b.u.template as<T>(), np, nz, nx, is_hio, ch, nh, nc);
  NUDIM_HIP(hipGetLastError());
  a.out(img);
}

extern "C" int nudim_constraint(int precision, double *img, const double *img_last, int nz, int np, int nx,
                                int is_hio, int config_height, int nheight, int ncenter) {
  API_BEGIN
  NUDIM_CHECK(img && img_last && nz > 0 && np > 0, "invalid argument");
  if (precision == NUDIM_FP32) constraint_parity<float>(img, img_last, nz, np, nx, is_hio, config_height, nheight, ncenter);
  else constraint_parity<double>(img, img_last, nz, np, nx, is_hio, config_height, nheight, ncenter);
  return NUDIM_OK;
  API_END
}

extern "C" int nudim_projection_ft(int precision, int nx, int nz, int np, int nmeas, const double *proj,
                                   double *tomo) {
  API_BEGIN
  (void)precision;  // set-up arithmetic is FP64 on both paths
  NUDIM_CHECK(proj && tomo && nx > 0 && nz > 0 && np > 0 && nmeas > 0, "invalid argument");
  TmpStream ts;
  size_t pb = (size_t)nmeas * nz * nx * 8, tb = (size_t)nmeas * nz * np * 16;
  DevBuf dp(pb), dt(tb);
  NUDIM_HIP(hipMemcpyAsync(dp.p, proj, pb, hipMemcpyHostToDevice, ts.s));
  projection_ft_dev(nx, nz, np, nmeas, dp.as<double>(), dt.as<cx<double>>(), ts.s);
  NUDIM_HIP(hipMemcpyAsync(tomo, dt.p, tb, hipMemcpyDeviceToHost, ts.s));
  NUDIM_HIP(hipStreamSynchronize(ts.s));
  return NUDIM_OK;
  API_END
}

// ---------------------------------------------------------------------------
// 2D1D forward simulator (projection_sim.c:282-405)
// ---------------------------------------------------------------------------

// density [z][y][x] (FP64) -> vol [y][x][z] with z taken at S(z) = (z+nz/2) mod nz:
// the z-FFT kernel reads vol[S(z)] = density[z], so its planes are the
// simulator's FFT along z (:319-322) at the shifted index (:336)
__global__ void k_sim_volume(const double *__restrict__ d, double *__restrict__ vol, int nx, int ny, int nz) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)nx * ny * nz) return;
  int z = (int)(i % nz);
  long long t = i / nz;
  int x = (int)(t % nx), y = (int)(t / nx);
  int zs = (z + nz / 2) % nz;
  vol[i] = d[((long long)zs * ny + y) * nx + x];
}

// c[(s*np + col)*nz + row] -> proj[s][row][col] = Re / (np * nz)  (:398-402)
__global__ void k_sim_out(const cx<double> *__restrict__ c, double *__restrict__ proj, int np, int nz, int ns) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)ns * nz * np) return;
  int col = (int)(i % np);
  long long t = i / np;
  int row = (int)(t % nz), s = (int)(t / nz);
  proj[i] = c[((long long)s * np + col) * nz + row].x / ((double)np * (double)nz);
}

// FFT along z, per-plane 2D NFFT (N = nx = ny, the reference's n = 2
// nextpow2(N) and 14 taps) at the tilt-range polar nodes, then per tilt the
// inverse 2D FFT of the (nz x np) Fourier slice with the reference's shifts:
// ifftshift of both axes going in (:362-367), fftshift of the columns coming
// out (:376-390).  FP64 throughout.
extern "C" int nudim_project_2d1d(int nx, int ny, int nz, int np, double gam_min, double gam_max, double interval,
                                  const double *density, double *proj) {
  API_BEGIN
  NUDIM_CHECK(density && proj, "invalid argument");
  NUDIM_CHECK(nx > 0 && ny > 0 && nz > 0 && np > 0 && interval > 0, "invalid dimensions or interval");
  NUDIM_CHECK(nx == ny, "the 2D1D simulator here takes square x-y planes (nx == ny)");
  NUDIM_CHECK(nx % 2 == 0 && nz % 2 == 0 && np % 2 == 0, "nx, nz and np must be even");
  TmpStream ts;
  const int N = nx, P = nz;
  std::vector<double> x;
  const int M = polar_nodes(gam_min, gam_max, interval, np, &x);
  const int ns = M / np;
  const long long V = (long long)nx * ny * nz;
  DevBuf dd(V * 8), vol(V * 8), fh(V * sizeof(cx<double>)), f0(V * sizeof(cx<double>));
  NUDIM_HIP(hipMemcpyAsync(dd.p, density, V * 8, hipMemcpyHostToDevice, ts.s));
  hipLaunchKernelGGL(k_sim_volume, dim3(cdiv(V, 256)), dim3(256), 0, ts.s, dd.as<double>(), vol.as<double>(), nx, ny,
                     nz);
  NUDIM_HIP(hipGetLastError());
  {
    // z FFT, planes innermost: fh[y][x][p] = FFT_z(density)[(p + nz/2) mod nz]
    FftDesc fdz = make_fft_desc(nz);
    std::vector<double> t;
    fft_twiddles(nz, &t);
    DevBuf twz((size_t)nz * sizeof(cx<double>));
    NUDIM_HIP(hipMemcpy(twz.p, t.data(), (size_t)nz * sizeof(cx<double>), hipMemcpyHostToDevice));
    int Gs = line_chunk_log2(nz, 1 << 30, (int)sizeof(cx<double>), FftEpt<double>::value);
    size_t lds = (size_t)nz * ((1 << Gs) + 1) * sizeof(cx<double>);
    long long nlines = (long long)nx * ny;
    hipLaunchKernelGGL(k_zfft_fwd<double>, dim3(cdiv(nlines, 1 << Gs)), dim3(kFftThreads), lds, ts.s,
                       vol.as<double>(), fh.as<cx<double>>(), f0.as<cx<double>>(), fdz, twz.as<cx<double>>(), Gs,
                       nlines, 1, 0LL);
    NUDIM_HIP(hipGetLastError());
  }
  NUDIM_HIP(hipStreamSynchronize(ts.s));
  dd.release();
  vol.release();
  f0.release();
  // per-plane NFFT: f[m][p], node m = s*np + c
  DevBuf f((size_t)M * P * sizeof(cx<double>)), g((size_t)M * P * sizeof(cx<double>));
  {
    NfftOp<double> op;
    op.init(N, 2 * next_pow2(N), 14, x.data(), M, P);
    op.trafo(fh.as<cx<double>>(), f.as<cx<double>>(), ts.s);
    NUDIM_HIP(hipStreamSynchronize(ts.s));
  }
  fh.release();
  // inverse FFT along z of each (s, c) line, input ifftshifted
  LineFft<double> fz, fc;
  fz.init(nz);
  fc.init(np);
  LineIO zi{nz, 1, nz, nz / 2, nullptr, nullptr}, zo{nz, 1, nz, 0, nullptr, nullptr};
  fz.run(f.as<cx<double>>(), g.as<cx<double>>(), M, zi, zo, +1, 1, ts.s);
  // inverse FFT along the columns of each tilt (rows = planes innermost),
  // input ifftshifted, output fftshifted
  const long long sso = (long long)np * nz;
  LineIO ci{sso, nz, np, np / 2, nullptr, nullptr}, co{sso, nz, np, np / 2, nullptr, nullptr};
  fc.run(g.as<cx<double>>(), f.as<cx<double>>(), ns, ci, co, +1, nz, ts.s);
  DevBuf out((size_t)ns * nz * np * 8);
  hipLaunchKernelGGL(k_sim_out, dim3(cdiv((long long)ns * nz * np, 256)), dim3(256), 0, ts.s, f.as<cx<double>>(),
                     out.as<double>(), np, nz, ns);
  NUDIM_HIP(hipGetLastError());
  NUDIM_HIP(hipMemcpyAsync(proj, out.p, (size_t)ns * nz * np * 8, hipMemcpyDeviceToHost, ts.s));
  NUDIM_HIP(hipStreamSynchronize(ts.s));
  return NUDIM_OK;
  API_END
}
