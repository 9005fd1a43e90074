// ops.hip -- batched NFFT, Toeplitz normal operator, per-plane reductions
// and the CGNR recurrence of the NFFT3 solver, for T = float and double.
#include "engine.hpp"

#include <algorithm>
#include <cmath>

namespace nudim {

thread_local KernelProfiler *g_prof = nullptr;

// ---------------------------------------------------------------------------
// small non-template kernels
// ---------------------------------------------------------------------------

__global__ void k_sum_partials(const double *__restrict__ part, int nb, int P, double *__restrict__ out) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  double s = 0;
  for (int b = 0; b < nb; b++) s += part[(size_t)b * P + p];
  out[p] = s;
}

__global__ void k_cg_scalars(double *__restrict__ dot_z, double *__restrict__ coef, const double *__restrict__ s,
                             int P, int mode) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  if (mode == 0) {
    dot_z[p] = s[p];
  } else if (mode == 1) {
    coef[p] = dot_z[p] / s[p];
  } else {
    double old = dot_z[p];
    dot_z[p] = s[p];
    coef[p] = s[p] / old;
  }
}

__global__ void k_sum_scalars(const double *__restrict__ part_g, const double *__restrict__ part_a, int nb,
                              double *__restrict__ sums) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double a = 0, b = 0;
    for (int i = 0; i < nb; i++) {
      a += part_g[i];
      b += part_a[i];
    }
    sums[0] = a;
    sums[1] = b;
  }
}

// ---------------------------------------------------------------------------
// FFT descriptors
// ---------------------------------------------------------------------------

FftDesc make_fft_desc(int n) {
  FftDesc d{};
  d.n = n;
  int m = n;
  NUDIM_CHECK(n >= 1, "FFT length must be positive");
  while (m % 8 == 0) { d.f[d.nf++] = 8; m /= 8; }
  while (m % 4 == 0) { d.f[d.nf++] = 4; m /= 4; }
  while (m % 2 == 0) { d.f[d.nf++] = 2; m /= 2; }
  while (m % 3 == 0) { d.f[d.nf++] = 3; m /= 3; }
  while (m % 5 == 0) { d.f[d.nf++] = 5; m /= 5; }
  NUDIM_CHECK(m == 1, "FFT length " + std::to_string(n) + " has a prime factor above 5");
  NUDIM_CHECK(d.nf <= 20, "FFT length too long");
  return d;
}

void fft_twiddles(int n, std::vector<double> *tw) {
  tw->resize(2 * (size_t)n);
  for (int j = 0; j < n; j++) {
    double a = -2.0 * M_PI * (double)j / (double)n;
    (*tw)[2 * j] = std::cos(a);
    (*tw)[2 * j + 1] = std::sin(a);
  }
}

template <typename T>
static void upload_cx(DevBuf &b, const std::vector<double> &v) {
  size_t cnt = v.size() / 2;
  std::vector<cx<T>> h(cnt);
  for (size_t i = 0; i < cnt; i++) h[i] = mk((T)v[2 * i], (T)v[2 * i + 1]);
  b.alloc(cnt * sizeof(cx<T>));
  NUDIM_HIP(hipMemcpy(b.p, h.data(), cnt * sizeof(cx<T>), hipMemcpyHostToDevice));
}

template <typename U>
static void upload(DevBuf &b, const std::vector<U> &v) {
  b.alloc(v.size() * sizeof(U));
  if (!v.empty()) NUDIM_HIP(hipMemcpy(b.p, v.data(), v.size() * sizeof(U), hipMemcpyHostToDevice));
}

int line_chunk_log2(int n, int P, int elem_bytes, int ept) {
  int cap = (kFftThreads * ept) / n;  // register stash bound
  int lds_cap = (64 * 1024) / (n * elem_bytes);
  cap = std::min(cap, lds_cap);
  cap = std::min(cap, 16);
  NUDIM_CHECK(cap >= 1, "line transform of length " + std::to_string(n) + " does not fit in LDS");
  int g = 0;
  while ((2 << g) <= cap && (P % (2 << g)) == 0) g++;
  return g;
}

template <typename T>
void LineFft<T>::init(int len) {
  n = len;
  fd = make_fft_desc(len);
  std::vector<double> t;
  fft_twiddles(len, &t);
  upload_cx<T>(tw, t);
}

template <typename T>
void LineFft<T>::run(const cx<T> *in, cx<T> *out, int nlines, const LineIO &li, const LineIO &lo, int sign,
                     int P, hipStream_t st, const cx<T> *mid, long long mid_so) const {
  int Gs = line_chunk_log2(n, P, (int)sizeof(cx<T>), FftEpt<T>::value);
  int nchunk = P >> Gs;
  size_t lds = (size_t)n * (1 << Gs) * sizeof(cx<T>);
  dim3 grid((unsigned)((long long)nlines * nchunk));
  double moved = (double)(li.cnt + lo.cnt) * nlines * P * sizeof(cx<T>) + (mid ? (double)n * nlines * sizeof(cx<T>) : 0.0);
  ProfScope ps(mid ? KC_TOEPLITZ_COL : KC_LINE_FFT, st, moved);
  hipLaunchKernelGGL(k_fft_lines<T>, grid, dim3(kFftThreads), lds, st, in, out, fd, li, lo, tw.as<cx<T>>(), sign, Gs,
                     nchunk, mid, mid_so);
  NUDIM_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// NFFT operator
// ---------------------------------------------------------------------------

template <typename T>
size_t NfftOp<T>::bytes() const {
  return fft.tw.bytes + c0.bytes + c1.bytes + psi0.bytes + psi1.bytes + dec.bytes + cell_ptr.bytes + ent.bytes +
         X.bytes + grid.bytes;
}

template <typename T>
void NfftOp<T>::init(int N_, int n_, int W_, const double *x, int M_, int P_) {
  N = N_; n = n_; W = W_; M = M_; P = P_;
  NUDIM_CHECK(N % 2 == 0 && n >= N && n % 2 == 0, "NFFT sizes must be even with n >= N");
  NUDIM_CHECK(W >= 2 && W <= 63 && W <= n, "window width out of range");
  NUDIM_CHECK(M < (1 << 20), "too many nodes for the gather map");
  NfftTables t;
  build_nfft_tables(N, n, W, x, M, true, &t);
  fft.init(n);
  upload(c0, t.c0);
  upload(c1, t.c1);
  upload(psi0, t.psi0);
  upload(psi1, t.psi1);
  upload(dec, t.dec);
  upload(cell_ptr, t.cell_ptr);
  upload(ent, t.ent);
  X.alloc((size_t)N * n * P * sizeof(cx<T>));
  grid.alloc((size_t)n * n * P * sizeof(cx<T>));
}

static int plane_group_log2(int P) {
  int s = 0;
  while ((2 << s) <= 64 && P % (2 << s) == 0) s++;
  return s;
}

template <typename T>
void NfftOp<T>::trafo(const cx<T> *fhat, cx<T> *f, hipStream_t st) {
  const long long NP = (long long)N * P, nP = (long long)n * P;
  const double *d = dec.as<double>();
  // rows: deconvolve, place at (k - N/2) mod n, FFT along x
  LineIO li{NP, P, N, -N / 2, d, d};
  LineIO lo{nP, P, n, 0, nullptr, nullptr};
  fft.run(fhat, X.as<cx<T>>(), N, li, lo, -1, P, st);
  // columns: N non-zero rows at (k - N/2) mod n, FFT along y
  LineIO ci{P, nP, N, -N / 2, nullptr, nullptr};
  LineIO co{P, nP, n, 0, nullptr, nullptr};
  fft.run(X.as<cx<T>>(), grid.as<cx<T>>(), n, ci, co, -1, P, st);
  int PCs = plane_group_log2(P);
  dim3 gr(cdiv(M, kThreads >> PCs), P >> PCs);
  // algorithmic bytes: grid read once, node values written, node coordinates
  ProfScope ps(KC_INTERP, st, ((double)n * n * sizeof(cx<T>) + (double)M * sizeof(cx<T>) + 2.0 * M * sizeof(T)) * P);
  hipLaunchKernelGGL(k_interp<T>, gr, dim3(kThreads), 0, st, grid.as<cx<T>>(), f, c0.as<int>(), c1.as<int>(),
                     psi0.as<double>(), psi1.as<double>(), M, n, W, P, PCs);
  NUDIM_HIP(hipGetLastError());
}

template <typename T>
void NfftOp<T>::adjoint(const cx<T> *f, cx<T> *fhat, hipStream_t st) {
  const long long NP = (long long)N * P, nP = (long long)n * P;
  int PCs = plane_group_log2(P);
  int ncell = n * n;
  dim3 gr(cdiv(ncell, kThreads >> PCs), P >> PCs);
  {
  // algorithmic bytes: node values read, node ids of the inverse map, grid written
  ProfScope ps(KC_GATHER, st, ((double)M * sizeof(cx<T>) + 4.0 * M + (double)ncell * sizeof(cx<T>)) * P);
  hipLaunchKernelGGL(k_gather<T>, gr, dim3(kThreads), 0, st, f, grid.as<cx<T>>(), cell_ptr.as<int>(),
                     ent.as<uint32_t>(), psi0.as<double>(), psi1.as<double>(), ncell, W, P, PCs);
  }
  NUDIM_HIP(hipGetLastError());
  const double *d = dec.as<double>();
  LineIO ci{P, nP, n, 0, nullptr, nullptr};
  LineIO co{P, nP, N, -N / 2, nullptr, nullptr};
  fft.run(grid.as<cx<T>>(), X.as<cx<T>>(), n, ci, co, +1, P, st);
  LineIO ri{nP, P, n, 0, nullptr, nullptr};
  LineIO ro{NP, P, N, -N / 2, d, d};
  fft.run(X.as<cx<T>>(), fhat, N, ri, ro, +1, P, st);
}

// ---------------------------------------------------------------------------
// Toeplitz normal operator
// ---------------------------------------------------------------------------

template <typename T>
__global__ void k_toeplitz_embed(const cx<double> *__restrict__ adj, cx<double> *__restrict__ emb, int N) {
  // adj: [2N][2N] adjoint values at frequency index k - N; emb: [L][L], L = 2N
  int L = 2 * N;
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)L * L) return;
  int r = (int)(i / L), c = (int)(i % L);
  int dy = r < N ? r : r - L, dx = c < N ? c : c - L;  // circulant index -> lag
  cx<double> v = mk(0.0, 0.0);
  if (dy > -N && dy < N && dx > -N && dx < N) v = adj[(long long)(dy + N) * L + (dx + N)];
  emb[i] = v;
}

template <typename T>
__global__ void k_toeplitz_finish(const cx<double> *__restrict__ spec, cx<T> *__restrict__ khat, int L, double scale) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)L * L) return;
  int r = (int)(i / L), c = (int)(i % L);
  cx<double> v = spec[i];
  khat[(long long)c * L + r] = mk((T)(v.x * scale), (T)(v.y * scale));
}

template <typename T>
size_t ToeplitzOp<T>::bytes() const {
  return fft.tw.bytes + khat.bytes + X.bytes;
}

template <typename T>
void ToeplitzOp<T>::init(int N_, const double *x, const double *w, int M, int P_, hipStream_t st) {
  N = N_;
  L = 2 * N;
  P = P_;
  fft.init(L);
  // Toeplitz kernel by one FP64 adjoint NFFT of the weights on a 2N grid
  // (reference-accuracy window: n = 2*nextpow2(2N), 14 taps)
  NfftOp<double> big;
  big.init(L, 2 * next_pow2(L), 14, x, M, 1);
  std::vector<double> wv(2 * (size_t)M);
  for (int m = 0; m < M; m++) {
    wv[2 * m] = w ? w[m] : 1.0;
    wv[2 * m + 1] = 0.0;
  }
  DevBuf dw;
  upload(dw, wv);
  DevBuf adj((size_t)L * L * sizeof(cx<double>)), emb((size_t)L * L * sizeof(cx<double>)),
      tmp((size_t)L * L * sizeof(cx<double>));
  big.adjoint(dw.as<cx<double>>(), adj.as<cx<double>>(), st);
  hipLaunchKernelGGL(k_toeplitz_embed<T>, dim3(cdiv((size_t)L * L, 256)), dim3(256), 0, st, adj.as<cx<double>>(),
                     emb.as<cx<double>>(), N);
  NUDIM_HIP(hipGetLastError());
  LineFft<double> f64;
  f64.init(L);
  LineIO ri{L, 1, L, 0, nullptr, nullptr};
  f64.run(emb.as<cx<double>>(), tmp.as<cx<double>>(), L, ri, ri, -1, 1, st);
  LineIO ci{1, L, L, 0, nullptr, nullptr};
  f64.run(tmp.as<cx<double>>(), emb.as<cx<double>>(), L, ci, ci, -1, 1, st);
  khat.alloc((size_t)L * L * sizeof(cx<T>));
  hipLaunchKernelGGL(k_toeplitz_finish<T>, dim3(cdiv((size_t)L * L, 256)), dim3(256), 0, st, emb.as<cx<double>>(),
                     khat.as<cx<T>>(), L, 1.0 / ((double)L * (double)L));
  NUDIM_HIP(hipGetLastError());
  NUDIM_HIP(hipStreamSynchronize(st));
  X.alloc((size_t)N * L * P * sizeof(cx<T>));
}

template <typename T>
void ToeplitzOp<T>::apply(const cx<T> *p, cx<T> *q, hipStream_t st) {
  const long long NP = (long long)N * P, LP = (long long)L * P;
  LineIO ri{NP, P, N, 0, nullptr, nullptr};
  LineIO ro{LP, P, L, 0, nullptr, nullptr};
  fft.run(p, X.as<cx<T>>(), N, ri, ro, -1, P, st);
  LineIO ci{P, LP, N, 0, nullptr, nullptr};
  fft.run(X.as<cx<T>>(), X.as<cx<T>>(), L, ci, ci, -1, P, st, khat.as<cx<T>>(), L);
  LineIO bi{LP, P, L, 0, nullptr, nullptr};
  LineIO bo{NP, P, N, 0, nullptr, nullptr};
  fft.run(X.as<cx<T>>(), q, N, bi, bo, +1, P, st);
}

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------

void Reducer::init(int P_) {
  P = P_;
  part.alloc((size_t)kRedBlocks * P * sizeof(double));
  part2.alloc((size_t)kRedBlocks * P * sizeof(double));
}

template <typename T>
void Reducer::dot(const cx<T> *a, const double *w, long long rows, double *out, hipStream_t st) {
  int PCs = plane_group_log2(P);
  dim3 gr(kRedBlocks, P >> PCs);
  hipLaunchKernelGGL((k_reduce<T, 0>), gr, dim3(kThreads), 0, st, a, (const cx<T> *)nullptr, w, (const int *)nullptr,
                     rows, P, PCs, part.as<double>(), (double *)nullptr);
  hipLaunchKernelGGL(k_sum_partials, dim3(cdiv(P, 64)), dim3(64), 0, st, part.as<double>(), kRedBlocks, P, out);
  NUDIM_HIP(hipGetLastError());
}

template <typename T>
void Reducer::redot(const cx<T> *a, const cx<T> *b, long long rows, double *out, hipStream_t st) {
  int PCs = plane_group_log2(P);
  dim3 gr(kRedBlocks, P >> PCs);
  hipLaunchKernelGGL((k_reduce<T, 2>), gr, dim3(kThreads), 0, st, a, b, (const double *)nullptr,
                     (const int *)nullptr, rows, P, PCs, part.as<double>(), (double *)nullptr);
  hipLaunchKernelGGL(k_sum_partials, dim3(cdiv(P, 64)), dim3(64), 0, st, part.as<double>(), kRedBlocks, P, out);
  NUDIM_HIP(hipGetLastError());
}

template <typename T>
void Reducer::rfactor(const cx<T> *f, const cx<T> *meas, const int *kind, long long rows, double *d1, double *d2,
                      hipStream_t st) {
  int PCs = plane_group_log2(P);
  dim3 gr(kRedBlocks, P >> PCs);
  hipLaunchKernelGGL((k_reduce<T, 1>), gr, dim3(kThreads), 0, st, f, meas, (const double *)nullptr, kind, rows, P,
                     PCs, part.as<double>(), part2.as<double>());
  hipLaunchKernelGGL(k_sum_partials, dim3(cdiv(P, 64)), dim3(64), 0, st, part.as<double>(), kRedBlocks, P, d1);
  hipLaunchKernelGGL(k_sum_partials, dim3(cdiv(P, 64)), dim3(64), 0, st, part2.as<double>(), kRedBlocks, P, d2);
  NUDIM_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// CGNR (NFFT3 solver order, CGNR branch)
// ---------------------------------------------------------------------------

template <typename T>
size_t Cgnr<T>::bytes() const {
  return z.bytes + pv.bytes + q.bytes + r.bytes + v.bytes + t.bytes + dot_z.bytes + coef.bytes + sc.bytes +
         red.part.bytes + red.part2.bytes;
}

template <typename T>
void Cgnr<T>::init(int N_, int M_, int P_, int form_, NfftOp<T> *op_, ToeplitzOp<T> *toep_, const double *w_) {
  N = N_; M = M_; P = P_; form = form_; op = op_; toep = toep_; w = w_;
  size_t NN = (size_t)N * N * P * sizeof(cx<T>), MM = (size_t)M * P * sizeof(cx<T>);
  z.alloc(NN);
  pv.alloc(NN);
  if (form == NUDIM_CG_TOEPLITZ) q.alloc(NN);
  r.alloc(MM);
  v.alloc(MM);
  t.alloc(MM);
  dot_z.alloc(P * sizeof(double));
  coef.alloc(P * sizeof(double));
  sc.alloc(P * sizeof(double));
  red.init(P);
}

template <typename T>
void Cgnr<T>::solve(const cx<T> *y, cx<T> *fhat, int nsteps, bool f0_zero, hipStream_t st) {
  const long long NN = (long long)N * N * P, MM = (long long)M * P;
  const unsigned bN = cdiv(NN, 256), bM = cdiv(MM, 256);
  cx<T> *zp = z.as<cx<T>>(), *pp = pv.as<cx<T>>(), *rp = r.as<cx<T>>(), *vp = v.as<cx<T>>(), *tp = t.as<cx<T>>();
  double *dz = dot_z.as<double>(), *cf = coef.as<double>(), *s = sc.as<double>();
  // before_loop: r = y - A f ; t = W r ; z = A^H t ; dot_z ; p = z
  if (!f0_zero) op->trafo(fhat, vp, st);
  hipLaunchKernelGGL(k_resid_init<T>, dim3(bM), dim3(256), 0, st, y, f0_zero ? (const cx<T> *)nullptr : vp, w, rp,
                     tp, MM, P);
  op->adjoint(tp, zp, st);
  red.dot<T>(zp, nullptr, (long long)N * N, s, st);
  hipLaunchKernelGGL(k_cg_scalars, dim3(cdiv(P, 64)), dim3(64), 0, st, dz, cf, s, P, 0);
  hipLaunchKernelGGL(k_copy_cx<T>, dim3(bN), dim3(256), 0, st, pp, zp, NN);
  NUDIM_HIP(hipGetLastError());
  for (int it = 0; it < nsteps; it++) {
    if (form == NUDIM_CG_TOEPLITZ) {
      // q = A^H W A p ; dot_v = <p, q> = sum w |A p|^2
      cx<T> *qp = q.as<cx<T>>();
      toep->apply(pp, qp, st);
      red.redot<T>(pp, qp, (long long)N * N, s, st);
      hipLaunchKernelGGL(k_cg_scalars, dim3(cdiv(P, 64)), dim3(64), 0, st, dz, cf, s, P, 1);
      hipLaunchKernelGGL(k_axpy_planes<T>, dim3(bN), dim3(256), 0, st, fhat, pp, cf, 1.0, NN, P);
      hipLaunchKernelGGL(k_axpy_planes<T>, dim3(bN), dim3(256), 0, st, zp, qp, cf, -1.0, NN, P);
    } else {
      // v = A p ; dot_v = sum w |v|^2
      op->trafo(pp, vp, st);
      red.dot<T>(vp, w, M, s, st);
      hipLaunchKernelGGL(k_cg_scalars, dim3(cdiv(P, 64)), dim3(64), 0, st, dz, cf, s, P, 1);
      hipLaunchKernelGGL(k_axpy_planes<T>, dim3(bN), dim3(256), 0, st, fhat, pp, cf, 1.0, NN, P);
      hipLaunchKernelGGL(k_resid_update<T>, dim3(bM), dim3(256), 0, st, rp, vp, cf, w, tp, MM, P);
      op->adjoint(tp, zp, st);
    }
    red.dot<T>(zp, nullptr, (long long)N * N, s, st);
    hipLaunchKernelGGL(k_cg_scalars, dim3(cdiv(P, 64)), dim3(64), 0, st, dz, cf, s, P, 2);
    hipLaunchKernelGGL(k_xpby_planes<T>, dim3(bN), dim3(256), 0, st, pp, zp, cf, NN, P);
    NUDIM_HIP(hipGetLastError());
  }
}

template struct LineFft<float>;
template struct LineFft<double>;
template struct NfftOp<float>;
template struct NfftOp<double>;
template struct ToeplitzOp<float>;
template struct ToeplitzOp<double>;
template struct Cgnr<float>;
template struct Cgnr<double>;
template void Reducer::dot<float>(const cx<float> *, const double *, long long, double *, hipStream_t);
template void Reducer::dot<double>(const cx<double> *, const double *, long long, double *, hipStream_t);
template void Reducer::redot<float>(const cx<float> *, const cx<float> *, long long, double *, hipStream_t);
template void Reducer::redot<double>(const cx<double> *, const cx<double> *, long long, double *, hipStream_t);
template void Reducer::rfactor<float>(const cx<float> *, const cx<float> *, const int *, long long, double *,
                                      double *, hipStream_t);
template void Reducer::rfactor<double>(const cx<double> *, const cx<double> *, const int *, long long, double *,
                                       double *, hipStream_t);

}  // namespace nudim
