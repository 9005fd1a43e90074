// ops.hip -- batched NFFT, Toeplitz normal operator, per-plane reductions
// and the CGNR recurrence of the NFFT3 solver, for T = float and double.
#include "engine.hpp"
#include "spread.hpp"

#include <algorithm>
#include <cstring>
#include <list>
#include <mutex>
#include <cmath>

namespace nudim {

thread_local KernelProfiler *g_prof = nullptr;
bool g_use_fft2 = getenv("NUDIM_GENERIC_FFT") == nullptr;
// patch-form spreading kernels (spread.hpp) where the sizes allow them;
// NUDIM_SPREAD=cell selects the per-node interpolation and per-cell gather
// the CG update z -= alpha q inside the Toeplitz row inverse pass, alpha
// from the column pass (NUDIM_CG_UPDATE=kernel: a kernel of its own after q)
static const bool g_cg_in_rowi = !(getenv("NUDIM_CG_UPDATE") && std::string(getenv("NUDIM_CG_UPDATE")) == "kernel");
static const bool g_spread_patch = !(getenv("NUDIM_SPREAD") && std::string(getenv("NUDIM_SPREAD")) == "cell");

// ---------------------------------------------------------------------------
// small non-template kernels
// ---------------------------------------------------------------------------

// block of 256 = 32 lanes x 8 planes; lane l sums rows l, l+32, ... and a
// fixed tree combines the lanes (same order on every run)
__global__ void k_sum_partials(const double *__restrict__ part, int nb, int P, double *__restrict__ out) {
  __shared__ double sh[256];
  const int pl = threadIdx.x & 7, lane = threadIdx.x >> 3;
  const int p = blockIdx.x * 8 + pl;
  double s = 0;
  // four strided partials in flight, added in index order
  if (p < P)
    for (int b0 = lane; b0 < nb; b0 += 32 * 4) {
      double v[4];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int b = b0 + 32 * k;
        v[k] = b < nb ? part[(size_t)b * P + p] : 0.0;
      }
#pragma unroll
      for (int k = 0; k < 4; k++) s += v[k];
    }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int h = 16; h > 0; h >>= 1) {
    if (lane < h) sh[threadIdx.x] += sh[threadIdx.x + h * 8];
    __syncthreads();
  }
  if (lane == 0 && p < P) out[p] = sh[pl];
}

// alpha[p] = zz[p] / (sum of the nb rows of part[row][stride] at plane p),
// p < np; the rows added as k_sum_partials adds them
__global__ void k_cg_alpha(const double *__restrict__ part, int nb, int stride, int np, const double *__restrict__ zz,
                           double *__restrict__ alpha) {
  __shared__ double sh[256];
  const int pl = threadIdx.x & 7, lane = threadIdx.x >> 3;
  const int p = blockIdx.x * 8 + pl;
  double s = 0;
  if (p < np)
    for (int b0 = lane; b0 < nb; b0 += 32 * 4) {
      double v[4];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int b = b0 + 32 * k;
        v[k] = b < nb ? part[(size_t)b * stride + p] : 0.0;
      }
#pragma unroll
      for (int k = 0; k < 4; k++) s += v[k];
    }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int h = 16; h > 0; h >>= 1) {
    if (lane < h) sh[threadIdx.x] += sh[threadIdx.x + h * 8];
    __syncthreads();
  }
  if (lane == 0 && p < np) alpha[p] = zz[p] / sh[pl];
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

// one block of 256 threads; fixed-order tree
__global__ void k_sum_scalars(const double *__restrict__ part_g, const double *__restrict__ part_a, int nb,
                              double *__restrict__ sums, int pst) {
  // volume blockIdx.x of a batch: its partial sums pst apart, its pair at sums + 2 blockIdx.x
  part_g += blockIdx.x * pst;
  part_a += blockIdx.x * pst;
  sums += 2 * blockIdx.x;
  __shared__ double sa[256], sb[256];
  double a = 0, b = 0;
  // eight strided values in flight, added in index order (the order of a
  // plain strided loop: adding 0 past the end changes nothing)
  for (int i0 = threadIdx.x; i0 < nb; i0 += 256 * 8) {
    double va[8], vb[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const int i = i0 + k * 256;
      va[k] = i < nb ? part_g[i] : 0.0;
      vb[k] = i < nb ? part_a[i] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < 8; k++) {
      a += va[k];
      b += vb[k];
    }
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
  if (threadIdx.x == 0) {
    sums[0] = sa[0];
    sums[1] = sb[0];
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
  int lds_cap = (160 * 1024) / (n * elem_bytes) - 1;  // lines + one twiddle table in 160 KB (gfx950)
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
  size_t lds = (size_t)n * ((1 << Gs) + 1) * sizeof(cx<T>);
  dim3 grid((unsigned)((long long)nlines * nchunk));
  double moved = (double)(li.cnt + lo.cnt) * nlines * P * sizeof(cx<T>) + (mid ? (double)n * nlines * sizeof(cx<T>) : 0.0);
  ProfScope ps(mid ? KC_TOEPLITZ_COL : KC_LINE_FFT, st, moved);
  if (g_use_fft2 && fft2_launch<T>(n, in, out, nlines, li, lo, sign, P, st, tw.as<cx<T>>(), mid, mid_so)) {
    NUDIM_HIP(hipGetLastError());
    return;
  }
  hipLaunchKernelGGL(k_fft_lines<T>, grid, dim3(kFftThreads), lds, st, in, out, fd, li, lo, tw.as<cx<T>>(), sign, Gs,
                     nchunk, mid, mid_so);
  NUDIM_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// NFFT operator
// ---------------------------------------------------------------------------

template <typename T>
size_t NfftGeom<T>::bytes() const {
  return fft.tw.bytes + c0.bytes + c1.bytes + psi0.bytes + psi1.bytes + dec.bytes + cell_ptr.bytes + ent.bytes + perm.bytes +
         gseg.bytes + gcomb.bytes + grec.bytes + iseg.bytes + irec.bytes + nptr.bytes + nent.bytes;
}

// Entry lists longer than smax are cut into nearly equal segments of at
// most that many entries, each its own work item (a wavefront or block walks
// one segment), so that no item walks the hundreds of entries of the
// patches where all rays meet.  seg: the segments, longest first (ties in
// patch order), empty patches last when kept; the segments of a split patch
// get partial-sum slots, and comb their (patch, first slot, count).
static void patch_segments(const PatchLists &pl, bool keep_empty, int smax, std::vector<PatchSeg> *seg,
                           std::vector<int> *comb, int *nslot) {
  seg->clear();
  comb->clear();
  int slot = 0;
  const int npatch = (int)pl.ptr.size() - 1;
  for (int q = 0; q < npatch; q++) {
    const int e0 = pl.ptr[q], cnt = pl.ptr[q + 1] - e0;
    if (cnt == 0) {
      if (keep_empty) seg->push_back(PatchSeg{q, e0, e0, -1});
      continue;
    }
    const int ns = (cnt + smax - 1) / smax;
    if (ns == 1) {
      seg->push_back(PatchSeg{q, e0, e0 + cnt, -1});
      continue;
    }
    comb->insert(comb->end(), {q, slot, ns});
    for (int k = 0; k < ns; k++)
      seg->push_back(PatchSeg{q, e0 + (int)((long long)k * cnt / ns), e0 + (int)((long long)(k + 1) * cnt / ns), slot++});
  }
  std::stable_sort(seg->begin(), seg->end(),
                   [](const PatchSeg &a, const PatchSeg &b) { return a.e1 - a.e0 > b.e1 - b.e0; });
  *nslot = slot;
}

// the gather's entry records, with one spare record at the end (RecChunk
// reads pieces past the last)
template <typename T, int R, int C>
static void gather_records(const PatchLists &pl, DevBuf *out) {
  std::vector<PatchRec<T, R, C>> gr(pl.ent.size() + 1);
  for (size_t e = 0; e < pl.ent.size(); e++) {
    const int m = (int)(pl.ent[e] & 0xFFFFFu), k0 = (int)((pl.ent[e] >> 20) & 63u), k1 = (int)(pl.ent[e] >> 26);
    const double *a0 = pl.psi0p.data() + (size_t)m * pl.l0 + k0, *a1 = pl.psi1p.data() + (size_t)m * pl.l1 + k1;
    gr[e].node = m;
    for (int i = 0; i < R; i++) gr[e].a0[i] = (T)a0[i];
    for (int j = 0; j < C; j++) gr[e].w1[j] = (T)a1[j];
  }
  upload(*out, gr);
}

template <typename T>
static std::shared_ptr<NfftGeom<T>> build_nfft_geometry(int N, int n, int W, const double *x, int M) {
  NUDIM_CHECK(N % 2 == 0 && n >= N && n % 2 == 0, "NFFT sizes must be even with n >= N");
  NUDIM_CHECK(W >= 2 && W <= 63 && W <= n, "window width out of range");
  NUDIM_CHECK(M < (1 << 20), "too many nodes for the gather map");
  auto g = std::make_shared<NfftGeom<T>>();
  g->N = N;
  g->n = n;
  g->W = W;
  g->M = M;
  NfftTables t;
  build_nfft_tables(N, n, W, x, M, true, &t);
  g->fft.init(n);
  upload(g->c0, t.c0);
  upload(g->c1, t.c1);
  upload(g->psi0, t.psi0);
  upload(g->psi1, t.psi1);
  upload(g->dec, t.dec);
  upload(g->cell_ptr, t.cell_ptr);
  upload(g->ent, t.ent);
  {
    auto wrap = [n](int c) { return ((c % n) + n) % n; };
    // Morton (Z) order of the 8 x 8-cell tiles, row-major inside a tile
    auto spread = [](unsigned v) {
      unsigned long long x = v & 0xFFFFu;
      x = (x | (x << 8)) & 0x00FF00FFull;
      x = (x | (x << 4)) & 0x0F0F0F0Full;
      x = (x | (x << 2)) & 0x33333333ull;
      x = (x | (x << 1)) & 0x55555555ull;
      return x;
    };
    std::vector<long long> key(M);
    for (int m = 0; m < M; m++) {
      const int r = wrap(t.c0[m]), c = wrap(t.c1[m]);
      const long long z = (long long)(spread((unsigned)(r / 8)) << 1 | spread((unsigned)(c / 8)));
      key[m] = (z * 8 + r % 8) * 8 + c % 8;
    }
    std::vector<int> perm(M);
    for (int m = 0; m < M; m++) perm[m] = m;
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return key[a] < key[b]; });
    upload(g->perm, perm);
  }
  {
    // gather: kGatherR x kGatherC-cell patches, one wavefront each
    PatchLists pl;
    g->has_patches = g_spread_patch && build_patch_lists(t, kGatherR, kGatherC, &pl);
    if (g->has_patches) {
      g->npatch = (int)pl.order.size();
      std::vector<PatchSeg> sg;
      std::vector<int> comb;
      // the gather's split patches cost a combine pass: segments of up to
      // 128 entries, or of the entries per wavefront slot of the GPU (3072)
      // when that is more, where the longest lists start first anyway
      patch_segments(pl, true, std::max<int>(128, (int)(pl.ent.size() / 3072)), &sg, &comb, &g->nslot);
      {
        // XCD x (blocks b with b % 8 = x) takes the patches of angular sector
        // x around the grid's origin (frequency 0 at cell 0, periodic), each
        // sector in the heavy-first order: position 8 k + x holds sector x's
        // k-th segment, the shorter sectors padded with pid = -1.  A node's
        // patches then meet in one XCD's L2 at about the same time.
        std::vector<PatchSeg> sec[8];
        const int npc = n / kGatherC;
        for (const PatchSeg &q : sg) {
          const int a = q.pid / npc, b = q.pid - a * npc;
          int cy = a * kGatherR + kGatherR / 2, cx = b * kGatherC + kGatherC / 2;
          if (cy >= n / 2) cy -= n;
          if (cx >= n / 2) cx -= n;
          const double ang = std::atan2((double)cy, (double)cx);
          const int xcd = std::min(7, std::max(0, (int)std::floor((ang + M_PI) / (2 * M_PI) * 8)));
          sec[xcd].push_back(q);
        }
        size_t len = 0;
        for (auto &v : sec) len = std::max(len, v.size());
        if (8 * len <= sg.size() + sg.size() / 2) {
          sg.assign(8 * len, PatchSeg{-1, 0, 0, -1});
          for (int xcd = 0; xcd < 8; xcd++)
            for (size_t k = 0; k < sec[xcd].size(); k++) sg[8 * k + xcd] = sec[xcd][k];
        } else {
          // node sets crowded into a few sectors (not the polar grids of a
          // job) keep the global heavy-first order, dealt round robin
          sg.resize((sg.size() + 7) / 8 * 8, PatchSeg{-1, 0, 0, -1});
        }
      }
      g->ngseg = (int)sg.size();
      g->ncomb = (int)comb.size() / 3;
      upload(g->gseg, sg);
      if (!comb.empty()) upload(g->gcomb, comb);
      gather_records<T, kGatherR, kGatherC>(pl, &g->grec);
    }
    // interpolation: 8 x 16-cell block patches, one block each; quarter q of
    // an entry is the 4 x 8 part at rows 4 (q / 2) .., columns 8 (q % 2) ..
    PatchLists bl;
    g->has_blocks = g->has_patches && build_patch_lists(t, 2 * kPatchR, 2 * kPatchC, &bl);
    if (g->has_blocks) {
      g->nbentry = (int)bl.ent.size();
      std::vector<PatchSeg> sg;
      std::vector<int> comb;
      int nslot = 0;
      patch_segments(bl, false, 128, &sg, &comb, &nslot);
      for (auto &q : sg) q.slot = -1;
      g->niseg = (int)sg.size();
      if (!sg.empty()) upload(g->iseg, sg);
      std::vector<PatchRec<T>> ir((bl.ent.size() + 1) * 4);
      for (size_t e = 0; e < bl.ent.size(); e++) {
        const int m = (int)(bl.ent[e] & 0xFFFFFu), k0 = (int)((bl.ent[e] >> 20) & 63u), k1 = (int)(bl.ent[e] >> 26);
        for (int q = 0; q < 4; q++) {
          // window taps of the quarter's first row and column
          const int t0 = k0 + kPatchR * (q >> 1) - (2 * kPatchR - 1), t1 = k1 + kPatchC * (q & 1) - (2 * kPatchC - 1);
          const bool meets = t0 + kPatchR > 0 && t0 < W && t1 + kPatchC > 0 && t1 < W;
          PatchRec<T> &r = ir[e * 4 + q];
          r.node = meets ? m : -1;
          const double *a0 = bl.psi0p.data() + (size_t)m * bl.l0 + k0 + kPatchR * (q >> 1);
          const double *a1 = bl.psi1p.data() + (size_t)m * bl.l1 + k1 + kPatchC * (q & 1);
          for (int i = 0; i < kPatchR; i++) r.a0[i] = (T)a0[i];
          for (int j = 0; j < kPatchC; j++) r.w1[j] = (T)a1[j];
        }
      }
      upload(g->irec, ir);
      upload(g->nptr, bl.node_ptr);
      upload(g->nent, bl.node_ent);
    }
  }
  return g;
}

namespace {
// most recently used geometries per precision (device ordinal, sizes, nodes)
template <typename T>
struct GeomCache {
  struct Entry {
    int dev, N, n, W, M;
    std::vector<double> x;
    std::shared_ptr<NfftGeom<T>> g;
  };
  std::mutex mu;
  std::list<Entry> lru;
};
// the caches keep up to kCacheEntries entries and kCacheBytes of device memory
constexpr size_t kCacheEntries = 64;
constexpr size_t kCacheBytes = (size_t)16 << 30;
}  // namespace

template <typename T>
std::shared_ptr<const NfftGeom<T>> nfft_geometry(int N, int n, int W, const double *x, int M) {
  // never destroyed: its device buffers must not be freed after the runtime
  // is torn down at exit
  static GeomCache<T> *cache = new GeomCache<T>();
  int dev = 0;
  NUDIM_HIP(hipGetDevice(&dev));
  auto same = [&](const typename GeomCache<T>::Entry &e) {
    return e.dev == dev && e.N == N && e.n == n && e.W == W && e.M == M &&
           std::memcmp(e.x.data(), x, sizeof(double) * 2 * (size_t)M) == 0;
  };
  {
    std::lock_guard<std::mutex> lk(cache->mu);
    for (auto it = cache->lru.begin(); it != cache->lru.end(); ++it)
      if (same(*it)) {
        cache->lru.splice(cache->lru.begin(), cache->lru, it);
        return cache->lru.front().g;
      }
  }
  // built outside the lock: two jobs of a new geometry may build it twice
  auto g = build_nfft_geometry<T>(N, n, W, x, M);
  std::lock_guard<std::mutex> lk(cache->mu);
  cache->lru.push_front({dev, N, n, W, M, std::vector<double>(x, x + 2 * (size_t)M), g});
  size_t total = 0;
  for (const auto &e : cache->lru) total += e.g->bytes();
  while (cache->lru.size() > 1 && (cache->lru.size() > kCacheEntries || total > kCacheBytes)) {
    total -= cache->lru.back().g->bytes();
    cache->lru.pop_back();
  }
  return g;
}

template <typename T>
size_t NfftOp<T>::bytes() const {
  return (g ? g->bytes() : 0) + part.bytes + gpart.bytes + X.bytes + grid.bytes;
}

template <typename T>
void NfftOp<T>::init(int N_, int n_, int W_, const double *x, int M_, int P_) {
  N = N_; n = n_; W = W_; M = M_; P = P_;
  g = nfft_geometry<T>(N, n, W, x, M);
  patched = g->has_patches && (P % 64 == 0);
  blocked = g->has_blocks && (P % 64 == 0);
  // the interpolation's partial sums are allocated by its first call: FP32
  // jobs run the Toeplitz operator and may never interpolate
  if (patched && g->nslot > 0) gpart.alloc((size_t)g->nslot * kGatherR * kGatherC * P * sizeof(cx<T>));
  X.alloc((size_t)N * n * P * sizeof(cx<T>));
  grid.alloc((size_t)n * n * P * sizeof(cx<T>));
}

// lanes per node / cell: the largest power of two <= cap dividing P
static int plane_group_log2(int P, int cap = 64) {
  int s = 0;
  while ((2 << s) <= cap && P % (2 << s) == 0) s++;
  return s;
}

// planes per cell group of the gather adjoint (a power of two up to 64):
// a whole wavefront on one cell's planes keeps every entry load coalesced
// across planes, which measured faster than several cells per wavefront
static const int g_gather_lanes = getenv("NUDIM_GATHER_LANES") ? atoi(getenv("NUDIM_GATHER_LANES")) : 64;

template <typename T>
void NfftOp<T>::grid_fft(const cx<T> *fhat, hipStream_t st) {
  const long long NP = (long long)N * P, nP = (long long)n * P;
  const double *d = g->dec.template as<double>();
  // the N columns of fhat first (deconvolved, placed at (k - N/2) mod n,
  // FFT along y): X[y][kx]; then the n rows, whose positions lie P elements
  // apart, write the grid (256^3 FP64: 2.68 -> 2.03 ms against columns
  // last, whose positions lie a grid row apart)
  LineIO ci{P, NP, N, -N / 2, d, d};
  LineIO co{P, NP, n, 0, nullptr, nullptr};
  g->fft.run(fhat, X.as<cx<T>>(), N, ci, co, -1, P, st);
  LineIO ri{NP, P, N, -N / 2, nullptr, nullptr};
  LineIO ro{nP, P, n, 0, nullptr, nullptr};
  g->fft.run(X.as<cx<T>>(), grid.as<cx<T>>(), n, ri, ro, -1, P, st);
}

template <typename T>
void NfftOp<T>::trafo_rsums(const cx<T> *fhat, const cx<T> *meas, const int *sel, int msel, double *part1,
                            double *part2, int nblocks, hipStream_t st) {
  grid_fft(fhat, st);
  const int PCs = plane_group_log2(P);
  const dim3 gr(nblocks, P >> PCs);
  // algorithmic bytes: grid read once, measured values and node coordinates
  ProfScope ps(KC_INTERP, st, ((double)n * n * sizeof(cx<T>) + (double)msel * sizeof(cx<T>) + 2.0 * msel * sizeof(T)) * P);
  const cx<T> *gp = grid.as<cx<T>>();
  const int *q0 = g->c0.template as<int>(), *q1 = g->c1.template as<int>();
  const double *p0 = g->psi0.template as<double>(), *p1 = g->psi1.template as<double>();
#define NUDIM_INTERP_R(WW)                                                                                        \
  case WW:                                                                                                        \
    hipLaunchKernelGGL((k_interp_rsums<T, WW>), gr, dim3(kThreads), 0, st, gp, meas, q0, q1, p0, p1, msel, n, P,  \
                       PCs, sel, part1, part2);                                                                   \
    break;
  switch (W) {
    NUDIM_INTERP_R(4)
    NUDIM_INTERP_R(6)
    NUDIM_INTERP_R(8)
    NUDIM_INTERP_R(10)
    NUDIM_INTERP_R(12)
    NUDIM_INTERP_R(14)
    NUDIM_INTERP_R(16)
    default:
      throw Error(NUDIM_EINVAL, "no fused R interpolation for window width " + std::to_string(W));
  }
#undef NUDIM_INTERP_R
  NUDIM_HIP(hipGetLastError());
}

template <typename T>
void NfftOp<T>::prepare_trafo() {
  if (blocked && !part.p) part.alloc((size_t)std::max(g->nbentry, 1) * P * sizeof(cx<T>));
}

template <typename T>
void NfftOp<T>::trafo(const cx<T> *fhat, cx<T> *f, hipStream_t st, const int *sel, int msel) {
  grid_fft(fhat, st);
  int PCs = plane_group_log2(P);
  const int Mc = sel ? msel : M;
  if (blocked && !sel) {
    prepare_trafo();
    // algorithmic bytes: grid read once, node values written, node coordinates
    ProfScope ps(KC_INTERP, st, ((double)n * n * sizeof(cx<T>) + (double)M * sizeof(cx<T>) + 2.0 * M * sizeof(T)) * P);
    const int G = P / 64;
    if (g->niseg > 0)
      hipLaunchKernelGGL(k_interp_block<T>, dim3((unsigned)((size_t)(g->niseg + 7) / 8 * 8 * G)), dim3(kThreads), 0, st,
                         grid.as<cx<T>>(), part.as<cx<T>>(), g->iseg.template as<PatchSeg>(),
                         g->irec.template as<PatchRec<T>>(), n, P, g->niseg);
    hipLaunchKernelGGL(k_interp_sum<T>, dim3(cdiv((size_t)M * G, 4)), dim3(kThreads), 0, st, part.as<cx<T>>(), f,
                       g->nptr.template as<int>(), g->nent.template as<int>(), M, P);
    NUDIM_HIP(hipGetLastError());
    return;
  }
  // every node: in the spatial order of the geometry (same values as node order)
  if (!sel && !(getenv("NUDIM_INTERP_ORDER") && std::string(getenv("NUDIM_INTERP_ORDER")) == "node"))
    sel = g->perm.template as<int>();
  dim3 gr(cdiv(Mc, kThreads >> PCs), P >> PCs);
  // algorithmic bytes: grid read once, node values written, node coordinates
  ProfScope ps(KC_INTERP, st, ((double)n * n * sizeof(cx<T>) + (double)Mc * sizeof(cx<T>) + 2.0 * Mc * sizeof(T)) * P);
  const cx<T> *gp = grid.as<cx<T>>();
  const int *q0 = g->c0.template as<int>(), *q1 = g->c1.template as<int>();
  const double *p0 = g->psi0.template as<double>(), *p1 = g->psi1.template as<double>();
#define NUDIM_INTERP_W(WW)                                                                                      \
  case WW:                                                                                                      \
    hipLaunchKernelGGL((k_interp_w<T, WW>), gr, dim3(kThreads), 0, st, gp, f, q0, q1, p0, p1, Mc, n, P, PCs, sel); \
    break;
  switch (W) {
    NUDIM_INTERP_W(4)
    NUDIM_INTERP_W(6)
    NUDIM_INTERP_W(8)
    NUDIM_INTERP_W(10)
    NUDIM_INTERP_W(12)
    NUDIM_INTERP_W(14)
    NUDIM_INTERP_W(16)
    default:
      hipLaunchKernelGGL(k_interp<T>, gr, dim3(kThreads), 0, st, gp, f, q0, q1, p0, p1, Mc, n, W, P, PCs, sel);
  }
#undef NUDIM_INTERP_W
  NUDIM_HIP(hipGetLastError());
}

template <typename T>
void NfftOp<T>::adjoint(const cx<T> *f, cx<T> *fhat, hipStream_t st) {
  const long long NP = (long long)N * P, nP = (long long)n * P;
  int PCs = plane_group_log2(P, g_gather_lanes);
  int ncell = n * n;
  dim3 gr(cdiv(ncell, kThreads >> PCs), P >> PCs);
  {
  // algorithmic bytes: node values read, node ids of the inverse map, grid written
  ProfScope ps(KC_GATHER, st, ((double)M * sizeof(cx<T>) + 4.0 * M + (double)ncell * sizeof(cx<T>)) * P);
  if (patched) {
    // 8 XCD sequences of ngseg / 8 items, G wavefronts each
    const unsigned gb = 8 * (unsigned)cdiv((size_t)g->ngseg / 8 * (P / 64), 4);
    hipLaunchKernelGGL((k_gather_patch<T, kGatherR, kGatherC>), dim3(gb), dim3(kThreads), 0, st, f, grid.as<cx<T>>(),
                       gpart.as<cx<T>>(), g->gseg.template as<PatchSeg>(),
                       g->grec.template as<PatchRec<T, kGatherR, kGatherC>>(), n, P, g->ngseg);
    if (g->ncomb > 0)
      hipLaunchKernelGGL((k_gather_combine<T, kGatherR, kGatherC>), dim3(cdiv((size_t)g->ncomb * kGatherR * (P / 64), 4)),
                         dim3(kThreads), 0, st, gpart.as<cx<T>>(), grid.as<cx<T>>(), g->gcomb.template as<int>(), n,
                         P, g->ncomb);
  } else {
    cx<T> *gp = grid.as<cx<T>>();
    const int *cp = g->cell_ptr.template as<int>();
    const uint32_t *ep = g->ent.template as<uint32_t>();
    const double *p0 = g->psi0.template as<double>(), *p1 = g->psi1.template as<double>();
#define NUDIM_GATHER_W(WW)                                                                                 \
  case WW:                                                                                                 \
    hipLaunchKernelGGL((k_gather_w<T, WW>), gr, dim3(kThreads), 0, st, f, gp, cp, ep, p0, p1, ncell, P, PCs); \
    break;
    switch (W) {
      NUDIM_GATHER_W(4)
      NUDIM_GATHER_W(6)
      NUDIM_GATHER_W(8)
      NUDIM_GATHER_W(10)
      NUDIM_GATHER_W(12)
      NUDIM_GATHER_W(14)
      NUDIM_GATHER_W(16)
      default:
        hipLaunchKernelGGL(k_gather<T>, gr, dim3(kThreads), 0, st, f, gp, cp, ep, p0, p1, ncell, W, P, PCs);
    }
#undef NUDIM_GATHER_W
  }
  }
  NUDIM_HIP(hipGetLastError());
  const double *d = g->dec.template as<double>();
  // the rows of the grid first (the n x n side, positions P elements apart;
  // its columns' positions lie a grid row apart, a slower access), N
  // outputs kept: X[y][kx]; then the N columns with the deconvolution
  LineIO ri{nP, P, n, 0, nullptr, nullptr};
  LineIO ro{NP, P, N, -N / 2, nullptr, nullptr};
  g->fft.run(grid.as<cx<T>>(), X.as<cx<T>>(), n, ri, ro, +1, P, st);
  LineIO ci{P, NP, n, 0, nullptr, nullptr};
  LineIO co{P, NP, N, -N / 2, d, d};
  g->fft.run(X.as<cx<T>>(), fhat, N, ci, co, +1, P, st);
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
__global__ void k_toeplitz_finish(const cx<double> *__restrict__ spec, cx<T> *__restrict__ khat, T *__restrict__ khat_re,
                                  int L, double scale) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)L * L) return;
  int r = (int)(i / L), c = (int)(i % L);
  cx<double> v = spec[i];
  if (khat) khat[(long long)c * L + r] = mk((T)(v.x * scale), (T)(v.y * scale));
  khat_re[(long long)c * L + r] = (T)(v.x * scale);
}

// Spectrum of the Toeplitz kernel T(d) = sum_m w_m exp(2 pi i d.x_m), d in
// [-N, N)^2, embedded in the 2N x 2N circulant, scaled by 1/L^2: one FP64
// adjoint NFFT of the weights on a 2N grid (reference-accuracy window:
// n = 2*nextpow2(2N), 14 taps), the embedding, an FP64 2D FFT.  khat (may be
// null) gets the complex spectrum, khat_re its real part.
template <typename T>
static void toeplitz_kernel(int N, const double *x, const double *w, int M, DevBuf *khat, DevBuf &khat_re,
                            hipStream_t st) {
  const int L = 2 * N;
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
  if (khat) khat->alloc((size_t)L * L * sizeof(cx<T>));
  khat_re.alloc((size_t)L * L * sizeof(T));
  hipLaunchKernelGGL(k_toeplitz_finish<T>, dim3(cdiv((size_t)L * L, 256)), dim3(256), 0, st, emb.as<cx<double>>(),
                     khat ? khat->as<cx<T>>() : nullptr, khat_re.as<T>(), L, 1.0 / ((double)L * (double)L));
  NUDIM_HIP(hipGetLastError());
  NUDIM_HIP(hipStreamSynchronize(st));
}

namespace {
template <typename T>
struct SpecCache {
  struct Entry {
    int dev, N, M;
    std::vector<double> x, w;  // w empty: unit weights
    std::shared_ptr<ToepSpectrum<T>> s;
  };
  std::mutex mu;
  std::list<Entry> lru;
};
}  // namespace

template <typename T>
std::shared_ptr<const ToepSpectrum<T>> toeplitz_spectrum(int N, const double *x, const double *w, int M,
                                                         hipStream_t st) {
  // never destroyed (device buffers must outlive any use at exit)
  static SpecCache<T> *cache = new SpecCache<T>();
  int dev = 0;
  NUDIM_HIP(hipGetDevice(&dev));
  auto same = [&](const typename SpecCache<T>::Entry &e) {
    return e.dev == dev && e.N == N && e.M == M && e.w.empty() == (w == nullptr) &&
           std::memcmp(e.x.data(), x, sizeof(double) * 2 * (size_t)M) == 0 &&
           (!w || std::memcmp(e.w.data(), w, sizeof(double) * (size_t)M) == 0);
  };
  {
    std::lock_guard<std::mutex> lk(cache->mu);
    for (auto it = cache->lru.begin(); it != cache->lru.end(); ++it)
      if (same(*it)) {
        cache->lru.splice(cache->lru.begin(), cache->lru, it);
        return cache->lru.front().s;
      }
  }
  auto sp = std::make_shared<ToepSpectrum<T>>();
  toeplitz_kernel<T>(N, x, w, M, &sp->khat, sp->khat_re, st);
  std::lock_guard<std::mutex> lk(cache->mu);
  cache->lru.push_front({dev, N, M, std::vector<double>(x, x + 2 * (size_t)M),
                         w ? std::vector<double>(w, w + M) : std::vector<double>(), sp});
  auto bytes = [](const ToepSpectrum<T> &t) { return t.khat.bytes + t.khat_re.bytes; };
  size_t total = 0;
  for (const auto &e : cache->lru) total += bytes(*e.s);
  while (cache->lru.size() > 1 && (cache->lru.size() > kCacheEntries || total > kCacheBytes)) {
    total -= bytes(*cache->lru.back().s);
    cache->lru.pop_back();
  }
  return sp;
}

template <typename T>
size_t ToeplitzOp<T>::bytes() const {
  return fft.tw.bytes + (spec ? spec->khat.bytes + spec->khat_re.bytes : 0) + X.bytes +
         (spec2 ? spec2->khat.bytes + spec2->khat_re.bytes : 0) + X2.bytes;
}

template <typename T>
void ToeplitzOp<T>::init_second(const double *x, const double *w2, int M, hipStream_t st) {
  spec2 = toeplitz_spectrum<T>(N, x, w2, M, st);
  X2.alloc((size_t)N * L * P * sizeof(cx<T>));
}

template <typename T>
void ToeplitzOp<T>::init_rform(const double *x, const double *w3, int M, hipStream_t st) {
  spec3 = toeplitz_spectrum<T>(N, x, w3, M, st);
}

template <typename T>
void ToeplitzOp<T>::apply_start(const cx<T> *a, const cx<T> *b, const cx<T> *add, cx<T> *z, double *zrows,
                                 hipStream_t st, double *qpart) {
  NUDIM_CHECK(fused() && spec2, "apply_start needs the fused passes and a second kernel");
  const long long NP = (long long)N * P, LC = (long long)L * pc;
  const cx<T> *tw = fft.tw.template as<cx<T>>();
  for (int p0 = 0; p0 < P; p0 += pc) {
    cx<T> *xa = X2.as<cx<T>>() + (long long)p0 * N * L, *xb = X.as<cx<T>>() + (long long)p0 * N * L;
    ToepCg<T> c{};
    // algorithmic bytes of a plain row pass: input read (N x N), X written (N x L)
    const double rowf_bytes = (double)(N + L) * N * pc * sizeof(cx<T>);
    {
      ProfScope ps(KC_TOEPLITZ_ROWF, st, rowf_bytes);
      toep_launch<T>(TK_ROWF, L, N, a + p0, xa, NP, P, LC, pc, pc, st, tw, nullptr, c, P, pairs);
    }
    {
      ProfScope ps(KC_TOEPLITZ_ROWF, st, rowf_bytes);
      toep_launch<T>(TK_ROWF, L, N, b + p0, xb, NP, P, LC, pc, pc, st, tw, nullptr, c, P, pairs);
    }
    {
      // X2 and X read, X written (N x L each), both kernels read
      ProfScope ps(KC_TOEPLITZ_COL2, st, 3.0 * N * L * pc * sizeof(cx<T>) + 2.0 * L * L * sizeof(T));
      c.in2 = xb;
      c.mid2 = spec->khat_re.template as<T>();
      if (qpart) {
        NUDIM_CHECK(spec3, "the quadratic form needs the third kernel (init_rform)");
        c.mid3 = spec3->khat_re.template as<T>();
        c.qpart = qpart + p0;
      }
      toep_launch<T>(TK_COL2, L, L, xa, xb, pc, LC, pc, LC, pc, st, tw, spec2->khat_re.template as<T>(), c, P, pairs);
    }
    {
      // X read (N x L), add read and z written (N x N each)
      ProfScope ps(KC_TOEPLITZ_ROWI, st, (2.0 * N + L) * N * pc * sizeof(cx<T>));
      ToepCg<T> ci{};
      ci.add = add + p0;
      if (zrows) ci.part = zrows + p0;
      toep_launch<T>(TK_ROWI, L, N, xb, z + p0, LC, pc, NP, P, pc, st, tw, nullptr, ci, P, pairs);
    }
  }
  NUDIM_HIP(hipGetLastError());
}

template <typename T>
void ToeplitzOp<T>::init(int N_, const double *x, const double *w, int M, int P_, hipStream_t st) {
  N = N_;
  L = 2 * N;
  P = P_;
  fft.init(L);
  spec = toeplitz_spectrum<T>(N, x, w, M, st);
  X.alloc((size_t)N * L * P * sizeof(cx<T>));
  // The column pass walks X's rows with stride L * pc: chunks of pc planes
  // (P halved while even and at least 64) keep a chunk's row within
  // NUDIM_TOEP_ROW_BYTES (default 2 MB).  Each plane's arithmetic is the
  // same for every chunking.
  const char *rb = getenv("NUDIM_TOEP_ROW_BYTES");
  const size_t row_cap = rb ? (size_t)strtoull(rb, nullptr, 10) : ((size_t)2 << 20);
  pc = P;
  while ((size_t)L * pc * sizeof(cx<T>) > row_cap && pc % 2 == 0 && pc / 2 >= 64) pc /= 2;
  const char *tp = getenv("NUDIM_TOEP_PAIRS");
  pairs = !(tp && std::string(tp) == "0");
}

template <typename T>
void ToeplitzOp<T>::apply(const cx<T> *p, cx<T> *q, hipStream_t st) {
  const long long NP = (long long)N * P, LP = (long long)L * P;
  LineIO ri{NP, P, N, 0, nullptr, nullptr};
  LineIO ro{LP, P, L, 0, nullptr, nullptr};
  fft.run(p, X.as<cx<T>>(), N, ri, ro, -1, P, st);
  LineIO ci{P, LP, N, 0, nullptr, nullptr};
  fft.run(X.as<cx<T>>(), X.as<cx<T>>(), L, ci, ci, -1, P, st, spec->khat.template as<cx<T>>(), L);
  LineIO bi{LP, P, L, 0, nullptr, nullptr};
  LineIO bo{NP, P, N, 0, nullptr, nullptr};
  fft.run(X.as<cx<T>>(), q, N, bi, bo, +1, P, st);
}

template <typename T>
void ToeplitzOp<T>::apply_cg(const ToepCg<T> &cg, cx<T> *q, hipStream_t st) {
  const long long NP = (long long)N * P, LC = (long long)L * pc;
  const cx<T> *tw = fft.tw.template as<cx<T>>();
  const T *kr = spec->khat_re.template as<T>();
  // one chunk of pc planes at a time; X of the chunk is [N][L][pc]
  for (int p0 = 0; p0 < P; p0 += pc) {
    cx<T> *x = X.as<cx<T>>() + (long long)p0 * N * L;
    ToepCg<T> c = cg;
    c.z += p0;
    c.p += p0;
    if (c.pin) c.pin += p0;
    if (c.f) c.f += p0;
    c.zz += p0;
    if (c.zz_old) c.zz_old += p0;
    if (c.alpha) c.alpha += p0;
    c.part += p0;
    {
      // algorithmic bytes: z read, p written, and with the recurrence p read,
      // f read and written (N x N each); X written (N x L)
      const double nvec = 2.0 + (cg.zz_old ? 1.0 : 0.0) + (cg.alpha ? 2.0 : 0.0);
      ProfScope ps(KC_TOEPLITZ_ROWF, st, (nvec * N + L) * N * pc * sizeof(cx<T>));
      toep_launch<T>(TK_ROWF, L, N, nullptr, x, NP, P, LC, pc, pc, st, tw, nullptr, c, P, pairs);
    }
    if (c.zu) {
      // <p, T p> from the column pass's spectrum, alpha, then the row
      // inverse pass updates z in place of writing q
      c.zu += p0;
      c.qpart += p0;
      {
        ProfScope ps(KC_TOEPLITZ_COL, st, 2.0 * N * L * pc * sizeof(cx<T>) + (double)L * L * sizeof(T));
        toep_launch<T>(TK_COL, L, L, x, x, pc, LC, pc, LC, pc, st, tw, kr, c, P, pairs);
      }
      hipLaunchKernelGGL(k_cg_alpha, dim3(cdiv(pc, 8)), dim3(256), 0, st, c.qpart, L, P, pc, c.zz,
                         const_cast<double *>(c.alz) + p0);
      c.alz += p0;
      {
        // algorithmic bytes: X read (N x L), z read and written (N x N each)
        ProfScope ps(KC_TOEPLITZ_ROWI, st, (2.0 * N + L) * N * pc * sizeof(cx<T>));
        ToepCg<T> ci{};
        ci.zu = c.zu;
        ci.alz = c.alz;
        ci.part = c.part;
        toep_launch<T>(TK_ROWI, L, N, x, c.zu, LC, pc, NP, P, pc, st, tw, nullptr, ci, P, pairs);
      }
      continue;
    }
    {
      ProfScope ps(KC_TOEPLITZ_COL, st, 2.0 * N * L * pc * sizeof(cx<T>) + (double)L * L * sizeof(T));
      toep_launch<T>(TK_COL, L, L, x, x, pc, LC, pc, LC, pc, st, tw, kr, c, P, pairs);
    }
    {
      // algorithmic bytes: X read (N x L), p read, q written (N x N each)
      ProfScope ps(KC_TOEPLITZ_ROWI, st, (2.0 * N + L) * N * pc * sizeof(cx<T>));
      toep_launch<T>(TK_ROWI, L, N, x, q + p0, LC, pc, NP, P, pc, st, tw, nullptr, c, P, pairs);
    }
  }
  NUDIM_HIP(hipGetLastError());
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
  hipLaunchKernelGGL(k_sum_partials, dim3(cdiv(P, 8)), dim3(256), 0, st, part.as<double>(), kRedBlocks, P, out);
  NUDIM_HIP(hipGetLastError());
}

template <typename T>
void Reducer::redot(const cx<T> *a, const cx<T> *b, long long rows, double *out, hipStream_t st) {
  int PCs = plane_group_log2(P);
  dim3 gr(kRedBlocks, P >> PCs);
  hipLaunchKernelGGL((k_reduce<T, 2>), gr, dim3(kThreads), 0, st, a, b, (const double *)nullptr,
                     (const int *)nullptr, rows, P, PCs, part.as<double>(), (double *)nullptr);
  hipLaunchKernelGGL(k_sum_partials, dim3(cdiv(P, 8)), dim3(256), 0, st, part.as<double>(), kRedBlocks, P, out);
  NUDIM_HIP(hipGetLastError());
}

template <typename T>
void Reducer::rfactor(const cx<T> *f, const cx<T> *meas, const int *kind, long long rows, double *d1, double *d2,
                      hipStream_t st) {
  int PCs = plane_group_log2(P);
  dim3 gr(kRedBlocks, P >> PCs);
  hipLaunchKernelGGL((k_reduce<T, 1>), gr, dim3(kThreads), 0, st, f, meas, (const double *)nullptr, kind, rows, P,
                     PCs, part.as<double>(), part2.as<double>());
  hipLaunchKernelGGL(k_sum_partials, dim3(cdiv(P, 8)), dim3(256), 0, st, part.as<double>(), kRedBlocks, P, d1);
  hipLaunchKernelGGL(k_sum_partials, dim3(cdiv(P, 8)), dim3(256), 0, st, part2.as<double>(), kRedBlocks, P, d2);
  NUDIM_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// CGNR (NFFT3 solver order, CGNR branch)
// ---------------------------------------------------------------------------

template <typename T>
size_t Cgnr<T>::bytes() const {
  return z.bytes + pv.bytes + q.bytes + r.bytes + v.bytes + t.bytes + dot_z.bytes + coef.bytes + sc.bytes +
         zz2.bytes + alpha.bytes + rowpart.bytes + colpart.bytes + pk.bytes + alk.bytes + red.part.bytes +
         red.part2.bytes;
}

template <typename T>
void Cgnr<T>::keep_directions(int K) {
  NUDIM_CHECK(form == NUDIM_CG_TOEPLITZ && toep->fused() && K > 0, "kept directions need the fused Toeplitz steps");
  pk.alloc((size_t)K * N * N * P * sizeof(cx<T>));
  alk.alloc((size_t)K * P * sizeof(double));
  kept = K;
}

template <typename T>
void Cgnr<T>::init(int N_, int M_, int P_, int form_, NfftOp<T> *op_, ToeplitzOp<T> *toep_, const double *w_) {
  N = N_; M = M_; P = P_; form = form_; op = op_; toep = toep_; w = w_;
  size_t NN = (size_t)N * N * P * sizeof(cx<T>), MM = (size_t)M * P * sizeof(cx<T>);
  z.alloc(NN);
  pv.alloc(NN);
  if (form == NUDIM_CG_TOEPLITZ) {
    q.alloc(NN);
    if (toep->fused()) {
      zz2.alloc(P * sizeof(double));
      alpha.alloc(P * sizeof(double));
      rowpart.alloc((size_t)N * P * sizeof(double));
      colpart.alloc((size_t)2 * N * P * sizeof(double));
    }
  }
  r.alloc(MM);
  v.alloc(MM);
  t.alloc(MM);
  dot_z.alloc(P * sizeof(double));
  coef.alloc(2 * P * sizeof(double));  // alpha, then beta of the NFFT-form step
  sc.alloc(P * sizeof(double));
  red.init(P);
}

template <typename T>
void Cgnr<T>::solve_from_residual(cx<T> *fhat, int nsteps, hipStream_t st, bool defer_last, bool z_rows) {
  NUDIM_CHECK(form == NUDIM_CG_TOEPLITZ && toep->fused(), "the fused Toeplitz passes are not available");
  const long long NN = (long long)N * N * P;
  const unsigned bN = cdiv(NN, 256);
  cx<T> *zp = z.as<cx<T>>(), *pp = pv.as<cx<T>>();
  double *dz = dot_z.as<double>(), *cf = coef.as<double>(), *s = sc.as<double>();
  // |z|^2 per plane: from the row sums the start's row pass left, or a read of z
  if (z_rows)
    hipLaunchKernelGGL(k_sum_partials, dim3(cdiv(P, 8)), dim3(256), 0, st, rowpart.as<double>(), N, P, s);
  else
    red.dot<T>(zp, nullptr, (long long)N * N, s, st);
  hipLaunchKernelGGL(k_cg_scalars, dim3(cdiv(P, 64)), dim3(64), 0, st, dz, cf, s, P, 0);
  NUDIM_HIP(hipGetLastError());
  // step k: row pass forms p_k = z_k + beta_k p_{k-1} (p_0 = z_0) and
  // applies the previous step's f += alpha p_{k-1} (kept directions: writes
  // p_k to its own buffer instead); q = T p_k with the row sums of
  // <p_k, q>; alpha_k = |z_k|^2 / <p_k, q>; z -= alpha_k q; |z_{k+1}|^2.
  // The last step's f update follows the loop.
  cx<T> *qp = q.as<cx<T>>();
  double *zz = dz, *zz_old = zz2.as<double>(), *pq = s, *al = alpha.as<double>();
  const int PCs = plane_group_log2(P);
  const dim3 rg(kRedBlocks, P >> PCs);
  const bool keep = defer_last && nsteps <= kept;
  for (int it = 0; it < nsteps; it++) {
    ToepCg<T> cg{zp, pp, fhat, zz, it ? zz_old : nullptr, it ? al : nullptr, rowpart.as<double>()};
    if (keep) {
      // p_k to its own buffer, p_{k-1} read from the previous one, no f update
      cg.p = pk.as<cx<T>>() + it * NN;
      cg.pin = it ? cg.p - NN : nullptr;
      cg.f = nullptr;
      cg.alpha = nullptr;
      al = alk.as<double>() + (long long)it * P;
    }
    if (g_cg_in_rowi) {
      // alpha from the column pass's <p, T p>; the row inverse pass takes
      // z -= alpha q and leaves the row sums of |z|^2 (q is never stored)
      cg.zu = zp;
      cg.alz = al;
      cg.qpart = colpart.as<double>();
      toep->apply_cg(cg, qp, st);
      hipLaunchKernelGGL(k_sum_partials, dim3(cdiv(P, 8)), dim3(256), 0, st, rowpart.as<double>(), N, P, zz_old);
      NUDIM_HIP(hipGetLastError());
      std::swap(zz, zz_old);
      continue;
    }
    toep->apply_cg(cg, qp, st);
    hipLaunchKernelGGL(k_sum_partials, dim3(cdiv(P, 8)), dim3(256), 0, st, rowpart.as<double>(), N, P, pq);
    {
      // algorithmic bytes: z, q read, z written
      ProfScope ps(KC_CG_UPDATE, st, 3.0 * NN * sizeof(cx<T>));
      hipLaunchKernelGGL(k_cg_zupdate<T>, rg, dim3(kThreads), 0, st, zp, qp, zz, pq, al, (long long)N * N, P, PCs,
                         red.part.as<double>());
    }
    hipLaunchKernelGGL(k_sum_partials, dim3(cdiv(P, 8)), dim3(256), 0, st, red.part.as<double>(), kRedBlocks, P,
                       zz_old);
    NUDIM_HIP(hipGetLastError());
    std::swap(zz, zz_old);
  }
  if (nsteps > 0 && !defer_last) {
    hipLaunchKernelGGL(k_axpy_planes<T>, dim3(bN), dim3(256), 0, st, fhat, pp, al, 1.0, NN, P);
    NUDIM_HIP(hipGetLastError());
  }
}

template <typename T>
void Cgnr<T>::solve(const cx<T> *y, cx<T> *fhat, int nsteps, bool f0_zero, hipStream_t st) {
  const long long NN = (long long)N * N * P, MM = (long long)M * P;
  const unsigned bN = cdiv(NN, 256), bM = cdiv(MM, 256);
  cx<T> *zp = z.as<cx<T>>(), *pp = pv.as<cx<T>>(), *rp = r.as<cx<T>>(), *vp = v.as<cx<T>>(), *tp = t.as<cx<T>>();
  double *dz = dot_z.as<double>(), *cf = coef.as<double>(), *s = sc.as<double>();
  const bool fused = form == NUDIM_CG_TOEPLITZ && toep->fused();
  if (fused) {
    // z = A^H W (y - A f) = A^H (W y) - T f: one adjoint and one normal-
    // operator pass (row pass with p := f, q = T f; p and the row sums are
    // overwritten by the first step)
    hipLaunchKernelGGL(k_resid_init<T>, dim3(bM), dim3(256), 0, st, y, (const cx<T> *)nullptr, w, rp, tp, MM, P);
    op->adjoint(tp, zp, st);
    if (!f0_zero) {
      cx<T> *qp = q.as<cx<T>>();
      ToepCg<T> c0{fhat, pp, nullptr, dz, nullptr, nullptr, rowpart.as<double>()};
      toep->apply_cg(c0, qp, st);
      hipLaunchKernelGGL(k_sub_cx<T>, dim3(bN), dim3(256), 0, st, zp, qp, NN);
    }
  } else {
    // before_loop: r = y - A f ; t = W r ; z = A^H t ; dot_z ; p = z
    if (!f0_zero) op->trafo(fhat, vp, st);
    hipLaunchKernelGGL(k_resid_init<T>, dim3(bM), dim3(256), 0, st, y, f0_zero ? (const cx<T> *)nullptr : vp, w, rp,
                       tp, MM, P);
    op->adjoint(tp, zp, st);
  }
  if (fused) {
    solve_from_residual(fhat, nsteps, st);
    return;
  }
  red.dot<T>(zp, nullptr, (long long)N * N, s, st);
  hipLaunchKernelGGL(k_cg_scalars, dim3(cdiv(P, 64)), dim3(64), 0, st, dz, cf, s, P, 0);
  NUDIM_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_copy_cx<T>, dim3(bN), dim3(256), 0, st, pp, zp, NN);
  NUDIM_HIP(hipGetLastError());
  if (form == NUDIM_CG_TOEPLITZ) {
    // q = A^H W A p;  <p, q> = sum w |A p|^2;  alpha = |z|^2 / <p, q>;
    // f += alpha p;  z -= alpha q;  beta = |z_new|^2 / |z|^2;  p = z + beta p
    cx<T> *qp = q.as<cx<T>>();
    double *zz = dz, *zz_new = cf, *pq = s;
    const int PCs = plane_group_log2(P);
    const dim3 rg(kRedBlocks, P >> PCs);
    for (int it = 0; it < nsteps; it++) {
      toep->apply(pp, qp, st);
      red.redot<T>(pp, qp, (long long)N * N, pq, st);
      hipLaunchKernelGGL(k_cg_update<T>, rg, dim3(kThreads), 0, st, fhat, pp, zp, qp, zz, pq, (long long)N * N, P,
                         PCs, red.part.as<double>());
      hipLaunchKernelGGL(k_sum_partials, dim3(cdiv(P, 8)), dim3(256), 0, st, red.part.as<double>(), kRedBlocks, P,
                         zz_new);
      hipLaunchKernelGGL(k_xpby_ratio<T>, dim3(bN), dim3(256), 0, st, pp, zp, zz_new, zz, NN, P);
      NUDIM_HIP(hipGetLastError());
      std::swap(zz, zz_new);
    }
    return;
  }
  for (int it = 0; it < nsteps; it++) {
    // v = A p ; dot_v = sum w |v|^2
    op->trafo(pp, vp, st);
    red.dot<T>(vp, w, M, s, st);
    hipLaunchKernelGGL(k_cg_scalars, dim3(cdiv(P, 64)), dim3(64), 0, st, dz, cf, s, P, 1);
    hipLaunchKernelGGL(k_resid_update<T>, dim3(bM), dim3(256), 0, st, rp, vp, cf, w, tp, MM, P);
    op->adjoint(tp, zp, st);
    red.dot<T>(zp, nullptr, (long long)N * N, s, st);
    // beta into the second half of coef; f += alpha p and p = z + beta p
    // of p's old value together (f is not read in between)
    hipLaunchKernelGGL(k_cg_scalars, dim3(cdiv(P, 64)), dim3(64), 0, st, dz, cf + P, s, P, 2);
    hipLaunchKernelGGL(k_fxp_planes<T>, dim3(bN), dim3(256), 0, st, fhat, pp, zp, cf, cf + P, NN, P);
    NUDIM_HIP(hipGetLastError());
  }
}

// ---------------------------------------------------------------------------
// TV sweeps (denoise.c:67-249): gradient pre-pass, niter-1 fused sweeps, and
// a final update, each followed by the fixed-order sum of the two norms
// ---------------------------------------------------------------------------

int tv_blocks(int ny, int nx, int nz) { return ny * 128; }

template <typename T>
T *tv_sweeps(T *u0, T *u1, T *g0, T *g1, int ny, int nx, int nz, int ylo, int yhi, float lam, int niter,
             double *part_g, double *part_a, double *sums, int nb, hipStream_t st, int K) {
  if (niter <= 0 || yhi < ylo) return u0;
  const long long rowlen = (long long)nx * nz;
  const long long V = (long long)ny * rowlen;
  const long long band = (long long)(yhi - ylo + 1) * rowlen;
  const double lamd = (double)lam;
  const double vb = (double)band * sizeof(T) * K;  // algorithmic: read u once + write u once per sweep
  // grid: (chunks of a row, band rows); one partial sum per block
  const int nrows = yhi - ylo + 1;
  const int gx = (int)std::min<long long>(cdiv(rowlen, kThreads), 64);
  const int np_ = nrows * gx;
  // sweeps on runs of 4 z per lane (k_tv_fused_q) when nz % 4 == 0.  (Runs
  // of 8, which make a wave cover a whole 512-long line, measured slower at
  // 512^3 in FP32: 583 against 534 us a sweep, at 3 waves per SIMD.)
  const bool quads = nz % 4 == 0;
  const int gx4 = quads ? (int)std::min<long long>(cdiv(rowlen / 4, kThreads), 128) : 0;
  // the same sweeps marching down y over chunks of R rows (k_tv_march)
  static const int march_rows = getenv("NUDIM_TV_ROWS") ? atoi(getenv("NUDIM_TV_ROWS")) : 8;
  const bool march = quads && march_rows > 0;
  const int R = std::max(1, march_rows);
  const dim3 mgrid(march ? (unsigned)cdiv(rowlen / 4, kThreads) : 1u, (unsigned)cdiv(nrows, R), K);
  const int npm = (int)(mgrid.x * mgrid.y);
  NUDIM_CHECK(np_ <= nb && nrows * gx4 <= nb && npm <= nb, "TV partial-sum buffer too small");
  NUDIM_CHECK((double)V * sizeof(T) < 4294967296.0, "TV kernels address volumes of fewer than 2^32 bytes");
  const dim3 grid(gx, nrows, K);
  // the sweeps write band rows only and read one row on each side of the
  // band, so u1 needs those two rows of u0; the last update writes into u0
  if (ylo > 0)
    hipLaunchKernelGGL(k_copy_real<T>, dim3(cdiv(rowlen, 256), K), dim3(256), 0, st, u1 + (ylo - 1) * rowlen,
                       u0 + (ylo - 1) * rowlen, rowlen, V);
  if (yhi < ny - 1)
    hipLaunchKernelGGL(k_copy_real<T>, dim3(cdiv(rowlen, 256), K), dim3(256), 0, st, u1 + (yhi + 1) * rowlen,
                       u0 + (yhi + 1) * rowlen, rowlen, V);
  {
    ProfScope ps(KC_TV, st, vb);
    hipLaunchKernelGGL(k_tv_grad<T>, grid, dim3(kThreads), 0, st, u0, g0, ny, nx, nz, ylo, yhi, part_g, part_a, V,
                       nb);
  }
  hipLaunchKernelGGL(k_sum_scalars, dim3(K), dim3(256), 0, st, part_g, part_a, np_, sums, nb);
  T *uc = u0, *un = u1, *gc = g0, *gn = g1;
  for (int s = 0; s + 1 < niter; s++) {
    {
      ProfScope ps(KC_TV, st, 2.0 * vb);
      if (march)
        hipLaunchKernelGGL(k_tv_march<T>, mgrid, dim3(kThreads), 0, st, uc, gc, un, gn, ny, nx, nz, ylo, yhi, R, sums,
                           lamd, part_g, part_a, V, nb);
      else if (quads)
        hipLaunchKernelGGL((k_tv_fused_q<T, 4>), dim3(gx4, nrows, K), dim3(kThreads), 0, st, uc, gc, un, gn, ny, nx,
                           nz, ylo, yhi, sums, lamd, part_g, part_a, V, nb);
      else
        hipLaunchKernelGGL(k_tv_fused<T>, grid, dim3(kThreads), 0, st, uc, gc, un, gn, ny, nx, nz, ylo, yhi, sums,
                           lamd, part_g, part_a, V, nb);
    }
    hipLaunchKernelGGL(k_sum_scalars, dim3(K), dim3(256), 0, st, part_g, part_a,
                       march ? npm : (quads ? nrows * gx4 : np_), sums, nb);
    std::swap(uc, un);
    std::swap(gc, gn);
  }
  {
    ProfScope ps(KC_TV, st, vb);
    hipLaunchKernelGGL(k_tv_update<T>, dim3(cdiv(band, 256), K), dim3(256), 0, st, u0, uc, gc, nx, nz, ylo, yhi, sums,
                       lamd, V);
  }
  NUDIM_HIP(hipGetLastError());
  return u0;
}

template <typename T>
void diffusion_sweeps(T *u, int ny, int nx, int nz, int niter, double lam, double ksq, hipStream_t st, int K) {
  if (niter <= 0) return;
  // algorithmic: read and write the volume once per sweep
  ProfScope ps(KC_DIFFUSION, st, 2.0 * niter * (double)ny * nx * nz * sizeof(T) * K);
  const size_t lds = (size_t)ny * nx * sizeof(T);
  const char *e = getenv("NUDIM_DIFFUSION_LDS");  // "0": the hyperplane launches only
  const bool lds_on = !(e && std::string(e) == "0");
  if (lds_on && lds <= (size_t)160 * 1024) {
    hipLaunchKernelGGL(k_diffusion_plane<T>, dim3(nz * K), dim3(1024), lds, st, u, ny, nx, nz, niter, lam, ksq);
    NUDIM_HIP(hipGetLastError());
    return;
  }
  const int taumax = 2 * (niter - 1) + (ny - 1) + (nx - 1);
  for (int k = 0; k < K; k++) {
    T *uk = u + (long long)k * ny * nx * nz;
    for (int tau = 0; tau <= taumax; tau++) {
      int smin = tau - (ny - 1) - (nx - 1);
      smin = smin > 0 ? (smin + 1) / 2 : 0;
      const int smax = std::min(niter - 1, tau / 2);
      if (smax < smin) continue;
      const long long cnt = (long long)(smax - smin + 1) * ny * nz;
      hipLaunchKernelGGL(k_diffusion_tau<T>, dim3(cdiv(cnt, 256)), dim3(256), 0, st, uk, ny, nx, nz, niter, tau,
                         lam, ksq);
    }
  }
  NUDIM_HIP(hipGetLastError());
}

template void diffusion_sweeps<float>(float *, int, int, int, int, double, double, hipStream_t, int);
template void diffusion_sweeps<double>(double *, int, int, int, int, double, double, hipStream_t, int);

template float *tv_sweeps<float>(float *, float *, float *, float *, int, int, int, int, int, float, int, double *,
                                   double *, double *, int, hipStream_t, int);
template double *tv_sweeps<double>(double *, double *, double *, double *, int, int, int, int, int, float, int,
                                     double *, double *, double *, int, hipStream_t, int);

template struct LineFft<float>;
template struct LineFft<double>;
template struct NfftGeom<float>;
template struct NfftGeom<double>;
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
