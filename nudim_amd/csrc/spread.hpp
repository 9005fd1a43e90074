// spread.hpp -- patch form of the NFFT window kernels (interpolation and
// gather adjoint) for plane counts that are multiples of 64.
//
// A wavefront owns a patch of grid cells (8 x 4 in the gather, 4 x 8 in the
// interpolation) and 64 planes (lane = plane), the patch held in registers,
// and walks the patch's entry list: the nodes whose window meets the patch,
// in node order (PatchLists in geom.hpp).  Every per-node quantity is
// wave-uniform -- the node and its window weights over the patch arrive 32
// (or 8) records at a time with vector loads and are broadcast with
// readlane -- and each node value is one coalesced 64-plane load per patch
// instead of one per tap.  The whole patch takes part in every entry:
// weights outside the node's window are zero.
//
// The gather writes every grid cell exactly once and needs no partial sums
// (patches with long lists are split into segments whose sums a second pass
// adds in segment order).
// The interpolation reads every grid cell once; a block of four wavefronts
// covers an 8 x 16-cell block patch, adds its quarters' sums through LDS and
// leaves one partial sum per (block patch, node) entry, which k_interp_sum
// adds per node in increasing block patch id.  Both agree with the per-cell
// gather and per-node interpolation (kernels.hpp) to rounding; zero weights
// times a non-finite grid value or node value can only reach cells or
// nodes of the same plane, which the 2D FFT beside them fills with
// non-finite values anyway.
#pragma once

#include "kernels.hpp"

#include <type_traits>

namespace nudim {

// patch shapes (rows x columns), 32 complex values per lane: the
// interpolation's quarter patches and the gather's patches (per entry, four
// columns of w1[j] s[node] and 32 products into the rows)
constexpr int kPatchR = 4, kPatchC = 8;
constexpr int kGatherR = 8, kGatherC = 4;

// One entry of a patch's list as the kernels read it: the node and its
// window weights over an R x C patch, a0[i] = psi0[node][t0] for patch row i
// and w1[j] = psi1[node][t1] for patch column j (zero outside the window), in
// the kernel's precision.
template <typename WT, int R = kPatchR, int C = kPatchC>
struct PatchRec {
  int node, pad[3];
  WT a0[R];
  WT w1[C];
};
static_assert(sizeof(PatchRec<double>) == 112 && sizeof(PatchRec<float>) == 64, "record layout");
static_assert(sizeof(PatchRec<double, kGatherR, kGatherC>) == 112 && sizeof(PatchRec<float, kGatherR, kGatherC>) == 64,
              "record layout");

// f(integral_constant<int, I>) for I = B .. E - 1, unrolled at compile time
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F &&f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// Wavefront q = 4 blockIdx.x + w (w = threadIdx.x / 64) takes item q / G of
// an order and planes 64 (q % G) ..; G = P / 64 plane groups.  With G = 4 a
// block is one patch over all planes: its four wavefronts read the same
// entry records and write whole cells (4 KB in FP64) together.
__device__ __forceinline__ void patch_wave(int G, int &item, int &p) {
  const int q = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
  item = q / G;
  p = (q - item * G) * 64 + (int)(threadIdx.x & 63);
}

// Up to K records (every stride-th from r) held in the wavefront's VGPRs:
// lane L keeps 16-byte pieces L / K + (64 / K) j (j < NJ) of record L % K,
// so dword d of record k sits in piece register (d / 4) / (64 / K), lane
// k + K ((d / 4) % (64 / K)), component d % 4.  Records reach the lanes with
// ordinary vector loads (a chunk costs one memory latency instead of one per
// entry) and each field is broadcast with readlane.  A record array carries
// one spare record at its end for the pieces read past the last record.
template <typename WT, int K, int R = kPatchR, int C = kPatchC>
struct RecChunk {
  using Rec = PatchRec<WT, R, C>;
  static constexpr int NP = (int)sizeof(Rec) / 16, NG = 64 / K, NJ = (NP + NG - 1) / NG;
  static_assert(sizeof(Rec) % 16 == 0 && NG * NJ - NP <= NP, "pieces past a record stay in the next");
  uint4 q[NJ];
  __device__ __forceinline__ void load(const Rec *__restrict__ r, int m, int stride) {
    const int lane = (int)(threadIdx.x & 63);
    const uint4 *rp = reinterpret_cast<const uint4 *>(r + (size_t)min(lane % K, m - 1) * stride) + lane / K;
#pragma unroll
    for (int j = 0; j < NJ; j++) q[j] = rp[NG * j];
  }
  template <int D>
  __device__ __forceinline__ unsigned dw(int k) const {
    const uint4 &v = q[(D / 4) / NG];
    const unsigned c = (D % 4) == 0 ? v.x : (D % 4) == 1 ? v.y : (D % 4) == 2 ? v.z : v.w;
    return (unsigned)__builtin_amdgcn_readlane((int)c, k + K * ((D / 4) % NG));
  }
  __device__ __forceinline__ int node(int k) const { return (int)dw<0>(k); }
  template <int D>
  __device__ __forceinline__ WT val(int k) const {
    if constexpr (sizeof(WT) == 8) {
      const unsigned long long lo = dw<D>(k), hi = dw<D + 1>(k);
      return __longlong_as_double((long long)(lo | (hi << 32)));
    } else {
      return __uint_as_float(dw<D>(k));
    }
  }
  template <int I>
  __device__ __forceinline__ WT a0(int k) const { return val<4 + I * (int)sizeof(WT) / 4>(k); }
  template <int J>
  __device__ __forceinline__ WT w1(int k) const { return val<4 + (R + J) * (int)sizeof(WT) / 4>(k); }
};

// one entry (record k of the chunk) added into the patch: column by column,
// w1[j] s[node] then a0[i] times that into each row (the weights outside the
// window are zero)
template <typename T, int R, int C>
__device__ __forceinline__ void gather_add(cx<T> (&acc)[R][C], const RecChunk<T, 32, R, C> &c, int k, cx<T> sv) {
  static_for<0, C>([&](auto J) {
    const cx<T> cs = c.template w1<J>(k) * sv;
    static_for<0, R>([&](auto I) { acc[I][J] = acc[I][J] + c.template a0<I>(k) * cs; });
  });
}

// One segment of a patch's entry list: the patch, entries [e0, e1) and the
// partial-sum slot (-1: the patch's only segment, which writes the grid).
struct PatchSeg {
  int pid, e0, e1, slot;
};

// g[cell][p] = sum over the cell's (node, t0, t1) of
//              psi0[node][t0] (psi1[node][t1] s[node][p]), in node order.
// A segment of a patch split into several (its entry list longer than one
// wavefront should walk alone) writes its sum to part[slot][cell][p] for
// k_gather_combine.
//
// The segments arrive dealt to the 8 XCDs: block b runs on XCD x = b % 8 and
// its wavefronts take the XCD's items il = 0, 1, .. (4 (b / 8) + w = G il +
// plane group), item 8 il + x of seg; seg places the segments of angular
// sector x of the grid (around its origin, heavy-first inside the sector) at
// the positions of XCD x, so that the patches sharing a node's value meet in
// one L2.  Padding items have pid < 0.
template <typename T, int PR, int PC>
__global__ void __launch_bounds__(kThreads)
k_gather_patch(const cx<T> *__restrict__ s, cx<T> *__restrict__ g, cx<T> *__restrict__ part,
               const PatchSeg *__restrict__ seg, const PatchRec<T, PR, PC> *__restrict__ rec, int n, int P,
               int nseg) {
  const int G = P >> 6, x = (int)(blockIdx.x & 7);
  const int lq = __builtin_amdgcn_readfirstlane((int)((blockIdx.x >> 3) * 4 + (threadIdx.x >> 6)));
  const int il = lq / G, item = il * 8 + x;
  const int p = (lq - il * G) * 64 + (int)(threadIdx.x & 63);
  if (item >= nseg) return;
  const PatchSeg sg = seg[item];
  if (sg.pid < 0) return;
  const int npc = n / PC, a = sg.pid / npc, b = sg.pid - a * npc;
  cx<T> acc[PR][PC];
#pragma unroll
  for (int i = 0; i < PR; i++)
#pragma unroll
    for (int j = 0; j < PC; j++) acc[i][j] = mk((T)0, (T)0);
  const int e0 = sg.e0, cnt = sg.e1 - sg.e0;
  for (int c0 = 0; c0 < cnt; c0 += 32) {
    const int m = min(32, cnt - c0);
    RecChunk<T, 32, PR, PC> ch;
    ch.load(rec + e0 + c0, m, 1);
    // three node values in flight: entries k, k + 1, k + 2 in sa, sb, sc;
    // each is reloaded (three entries ahead, clamped to the chunk's last)
    // after its use
    cx<T> sa = s[(size_t)ch.node(0) * P + p];
    cx<T> sb = s[(size_t)ch.node(min(1, m - 1)) * P + p];
    cx<T> sc = s[(size_t)ch.node(min(2, m - 1)) * P + p];
    int k = 0;
    for (; k + 3 <= m; k += 3) {
      gather_add(acc, ch, k, sa);
      sa = s[(size_t)ch.node(min(k + 3, m - 1)) * P + p];
      gather_add(acc, ch, k + 1, sb);
      sb = s[(size_t)ch.node(min(k + 4, m - 1)) * P + p];
      gather_add(acc, ch, k + 2, sc);
      sc = s[(size_t)ch.node(min(k + 5, m - 1)) * P + p];
    }
    if (k < m) gather_add(acc, ch, k, sa);
    if (k + 1 < m) gather_add(acc, ch, k + 1, sb);
  }
  if (sg.slot < 0) {
#pragma unroll
    for (int i = 0; i < PR; i++)
#pragma unroll
      for (int j = 0; j < PC; j++) g[((size_t)(a * PR + i) * n + (size_t)(b * PC + j)) * P + p] = acc[i][j];
  } else {
#pragma unroll
    for (int i = 0; i < PR; i++)
#pragma unroll
      for (int j = 0; j < PC; j++) part[((size_t)sg.slot * PR * PC + i * PC + j) * P + p] = acc[i][j];
  }
}

// The patches split into segments: cell = the segments' partial sums added
// in segment order.  comb holds (patch, first slot, count) triples; one
// wavefront per (patch, patch row) and plane group (item = triple PR + row,
// the wavefront layout of patch_wave), two segments' loads in flight.
template <typename T, int PR, int PC>
__global__ void __launch_bounds__(kThreads)
k_gather_combine(const cx<T> *__restrict__ part, cx<T> *__restrict__ g, const int *__restrict__ comb, int n, int P,
                 int ncomb) {
  int item, p;
  patch_wave(P >> 6, item, p);
  if (item >= ncomb * PR) return;
  const int c = item / PR, i = item - c * PR;
  const int pid = comb[3 * c], slot0 = comb[3 * c + 1], nseg = comb[3 * c + 2];
  const int npc = n / PC, a = pid / npc, b = pid - a * npc;
  const cx<T> *src = part + ((size_t)slot0 * PR * PC + i * PC) * P + p;
  const size_t ss = (size_t)PR * PC * P;  // one segment's partial sums
  cx<T> acc[PC];
#pragma unroll
  for (int j = 0; j < PC; j++) acc[j] = src[(size_t)j * P];
  int k = 1;
  for (; k + 2 <= nseg; k += 2) {
    cx<T> v[PC], w[PC];
#pragma unroll
    for (int j = 0; j < PC; j++) v[j] = src[k * ss + (size_t)j * P];
#pragma unroll
    for (int j = 0; j < PC; j++) w[j] = src[(k + 1) * ss + (size_t)j * P];
#pragma unroll
    for (int j = 0; j < PC; j++) acc[j] = (acc[j] + v[j]) + w[j];
  }
  if (k < nseg) {
#pragma unroll
    for (int j = 0; j < PC; j++) acc[j] = acc[j] + src[k * ss + (size_t)j * P];
  }
#pragma unroll
  for (int j = 0; j < PC; j++) g[((size_t)(a * PR + i) * n + (size_t)(b * PC + j)) * P + p] = acc[j];
}

template <int I, typename T>
__device__ __forceinline__ void interp_row(cx<T> &acc, const cx<T> (&v)[kPatchR][kPatchC], const RecChunk<T, 8> &c,
                                           int k) {
  cx<T> racc = mk((T)0, (T)0);
#define NUDIM_ICOL(J) racc = racc + c.template w1<J>(k) * v[I][J];
  NUDIM_ICOL(0) NUDIM_ICOL(1) NUDIM_ICOL(2) NUDIM_ICOL(3) NUDIM_ICOL(4) NUDIM_ICOL(5) NUDIM_ICOL(6) NUDIM_ICOL(7)
#undef NUDIM_ICOL
  acc = acc + c.template a0<I>(k) * racc;
}

// Interpolation, first half.  A block takes one segment (a range of at most
// a few hundred entries) of an 8 x 16-cell block patch and 64 planes.
// Wavefront w holds the 4 x 8 quarter (rows 4 (w / 2) ..,
// columns 8 (w % 2) ..) in registers and reads quarter-record w of each
// entry (node < 0: the quarter is outside the node's window).  Per entry it
// forms sum over the quarter's rows of a0[i] (sum over the row's cells of
// w1[j] g[cell][p]) -- the row-then-column order of k_interp_w on its share
// of the window -- and the four quarters' sums, added in quarter order
// through LDS eight entries at a time, give part[e][p].
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_interp_block(const cx<T> *__restrict__ g, cx<T> *__restrict__ part, const PatchSeg *__restrict__ seg,
               const PatchRec<T> *__restrict__ rec, int n, int P, int nitem) {
  constexpr int PR = kPatchR, PC = kPatchC, KB = 8;
  __shared__ cx<T> red[KB][4][64];
  // block b runs on XCD b % 8 and is its (b / 8)-th block there: XCD x takes
  // segments x, x + 8, .. of the order, each over its G plane groups in turn
  // (the blocks reading one segment's records share an L2)
  const int G = P >> 6, i8 = (int)blockIdx.x >> 3;
  const int item = (i8 / G) * 8 + (int)(blockIdx.x & 7);
  if (item >= nitem) return;
  const int lane = (int)(threadIdx.x & 63);
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int p = (i8 % G) * 64 + lane;
  const PatchSeg sg = seg[item];
  const int pid = sg.pid;
  const int nbc = n / (2 * PC), a = pid / nbc, b = pid - a * nbc;
  const int r0 = a * 2 * PR + (w >> 1) * PR, q0 = b * 2 * PC + (w & 1) * PC;
  cx<T> v[PR][PC];
#pragma unroll
  for (int i = 0; i < PR; i++)
#pragma unroll
    for (int j = 0; j < PC; j++) v[i][j] = g[((size_t)(r0 + i) * n + (size_t)(q0 + j)) * P + p];
  const int e0 = sg.e0, cnt = sg.e1 - sg.e0;
  for (int c0 = 0; c0 < cnt; c0 += KB) {
    const int m = min(KB, cnt - c0);
    RecChunk<T, KB> ch;
    ch.load(rec + (size_t)(e0 + c0) * 4 + w, m, 4);
    for (int k = 0; k < m; k++) {
      cx<T> acc = mk((T)0, (T)0);
      if (ch.node(k) >= 0) {
        interp_row<0>(acc, v, ch, k);
        interp_row<1>(acc, v, ch, k);
        interp_row<2>(acc, v, ch, k);
        interp_row<3>(acc, v, ch, k);
      }
      red[k][w][lane] = acc;
    }
    __syncthreads();
    for (int k = w; k < m; k += 4)
      part[(size_t)(e0 + c0 + k) * P + p] = ((red[k][0][lane] + red[k][1][lane]) + red[k][2][lane]) + red[k][3][lane];
    __syncthreads();
  }
}

// Interpolation, second half: out[node][p] = the node's partial sums added in
// increasing block patch id.  One wavefront per node and 64 planes (as
// patch_wave).
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_interp_sum(const cx<T> *__restrict__ part, cx<T> *__restrict__ out, const int *__restrict__ nptr,
             const int *__restrict__ nent, int M, int P) {
  int node, p;
  patch_wave(P >> 6, node, p);
  if (node >= M) return;
  const int k0 = nptr[node], k1 = nptr[node + 1];
  cx<T> acc = mk((T)0, (T)0);
  int k = k0;
  // four partial sums in flight, added in order
  for (; k + 4 <= k1; k += 4) {
    cx<T> q[4];
#pragma unroll
    for (int u = 0; u < 4; u++) q[u] = part[(size_t)nent[k + u] * P + p];
#pragma unroll
    for (int u = 0; u < 4; u++) acc = acc + q[u];
  }
  for (; k < k1; k++) acc = acc + part[(size_t)nent[k] * P + p];
  out[(size_t)node * P + p] = acc;
}

}  // namespace nudim
