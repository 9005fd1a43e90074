// geom.cpp -- host-side set-up tables (see geom.hpp).
#include "geom.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <stdexcept>

namespace nudim {

int polar_nodes(double amin, double amax, double ainterval, int np, std::vector<double> *x) {
  int nslices = (int)((amax - amin) / ainterval + 0.5) + 1;
  if (x) {
    x->assign((size_t)2 * nslices * np, 0.0);
    double angle = amin - ainterval;
    for (int s = 0; s < nslices; s++) {
      angle += ainterval;
      double interval = 1.0 / np;
      double r = -0.5 - interval;
      double *xs = x->data() + (size_t)s * np * 2;
      for (int i = 0; i < np; i++) {
        r += interval;
        xs[2 * i + 0] = r * std::sin(angle / 180 * kRefPi);
        xs[2 * i + 1] = r * std::cos(angle / 180 * kRefPi);
      }
    }
  }
  return nslices * np;
}

void uniform_nodes(int nx, int nz, std::vector<double> *x) {
  x->assign((size_t)2 * nx * nz, 0.0);
  double ivx = 1.0 / nx, ivz = 1.0 / nz;
  double z0 = -0.5 - ivz;
  for (int j = 0; j < nz; j++) {
    z0 += ivz;
    double x0 = -0.5 - ivx;
    for (int i = 0; i < nx; i++) {
      x0 += ivx;
      (*x)[2 * ((size_t)j * nx + i) + 0] = x0;
      (*x)[2 * ((size_t)j * nx + i) + 1] = z0;
    }
  }
}

void pad_range(double gam_min, double interval, double *pad_min, double *pad_max) {
  int t = (int)((-90 - gam_min) / interval);
  *pad_min = gam_min + t * interval;
  t = (int)((90 - gam_min) / interval);
  *pad_max = gam_min + t * interval;
}

void density_weights(int np, int nang, double padangle, std::vector<double> *w) {
  w->assign((size_t)nang * np, 0.0);
  for (int j = 0; j < nang; j++)
    for (int i = 0; i < np; i++) {
      double v;
      if (i == np / 2)
        v = kRefPi / (4 * nang * std::pow(np / 2, 2));
      else
        v = padangle / 180 * kRefPi * std::abs(i - np / 2) / std::pow(np / 2, 2);
      (*w)[(size_t)j * np + i] = v;
    }
}

void sample_classes(int np, int nang, double pad_min, double padangle, double gam_min,
                    double gam_max, double interval, std::vector<int> *kind,
                    std::vector<int> *order) {
  int M = nang * np;
  kind->assign(M, 0);
  order->assign(M, -1);
  for (int i = 0; i < M; i++) {
    int nslc = i / np;
    double angle = pad_min + nslc * padangle;
    int span = (int)(interval / padangle + 0.5);
    int seq = (int)((angle - gam_min) / padangle + 0.5);
    if ((angle < gam_min) || (angle > gam_max) || (seq % span != 0)) {
      (*kind)[i] = 0;
    } else if (std::abs((int)angle) < 1e-5) {
      (*kind)[i] = 2;
    } else {
      (*kind)[i] = 1;
      (*order)[i] = (int)((angle - gam_min) / interval + 0.5);
    }
  }
}

double bessel_i0(double x) {
  double q = 0.25 * x * x, term = 1.0, sum = 1.0;
  for (int k = 1; k < 500; k++) {
    term *= q / ((double)k * (double)k);
    sum += term;
    if (term < 1e-17 * sum) break;
  }
  return sum;
}

double kb_beta(int W, double sigma) {
  double a = (double)W / sigma * (sigma - 0.5);
  return M_PI * std::sqrt(a * a - 0.8);
}

double kb_psi(double d, int W, double beta) {
  double t = 2.0 * d / W, s = 1.0 - t * t;
  if (s < 0) return 0.0;
  return bessel_i0(beta * std::sqrt(s)) / bessel_i0(beta);
}

// Fourier transform of kb_psi: W sinh(sqrt(b^2 - (pi W nu)^2)) / sqrt(...) / I0(b)
double kb_psihat(double nu, int W, double beta) {
  double a = M_PI * W * nu, s = beta * beta - a * a, v;
  if (s > 1e-300) {
    double r = std::sqrt(s);
    v = std::sinh(r) / r;
  } else if (s < -1e-300) {
    double r = std::sqrt(-s);
    v = std::sin(r) / r;
  } else {
    v = 1.0;
  }
  return W * v / bessel_i0(beta);
}

static inline int pmod(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}

void build_nfft_tables(int N, int n, int W, const double *x, int M, bool with_gather,
                       NfftTables *t) {
  t->N = N; t->n = n; t->W = W; t->M = M;
  double beta = kb_beta(W, (double)n / N);
  double i0b = bessel_i0(beta);
  t->c0.resize(M); t->c1.resize(M);
  t->psi0.resize((size_t)M * W); t->psi1.resize((size_t)M * W);
  for (int m = 0; m < M; m++) {
    for (int d = 0; d < 2; d++) {
      double u = x[2 * m + d] * n;
      int c = (int)std::floor(u) - W / 2 + 1;
      double *psi = (d == 0 ? t->psi0.data() : t->psi1.data()) + (size_t)m * W;
      for (int k = 0; k < W; k++) {
        double dd = u - (double)(c + k), tt = 2.0 * dd / W, s = 1.0 - tt * tt;
        psi[k] = s < 0 ? 0.0 : bessel_i0(beta * std::sqrt(s)) / i0b;
      }
      (d == 0 ? t->c0 : t->c1)[m] = c;
    }
  }
  t->dec.resize(N);
  for (int k = 0; k < N; k++) t->dec[k] = 1.0 / kb_psihat((double)(k - N / 2) / n, W, beta);
  t->cell_ptr.clear();
  t->ent.clear();
  if (!with_gather) return;
  size_t ncell = (size_t)n * n;
  std::vector<int> cnt(ncell + 1, 0);
  for (int m = 0; m < M; m++)
    for (int a = 0; a < W; a++) {
      int r = pmod(t->c0[m] + a, n);
      for (int b = 0; b < W; b++) cnt[(size_t)r * n + pmod(t->c1[m] + b, n) + 1]++;
    }
  for (size_t c = 0; c < ncell; c++) cnt[c + 1] += cnt[c];
  t->cell_ptr.assign(cnt.begin(), cnt.end());
  t->ent.resize((size_t)cnt[ncell]);
  std::vector<int> fill(cnt.begin(), cnt.end() - 1);
  for (int m = 0; m < M; m++)
    for (int a = 0; a < W; a++) {
      int r = pmod(t->c0[m] + a, n);
      for (int b = 0; b < W; b++) {
        size_t c = (size_t)r * n + pmod(t->c1[m] + b, n);
        t->ent[fill[c]++] = (uint32_t)m | ((uint32_t)a << 20) | ((uint32_t)b << 26);
      }
    }
}

void build_gather_tiles(const NfftTables &t, int ts, int k, GatherTiles *g) {
  const int n = t.n, W = t.W, M = t.M;
  if (n % ts != 0) throw std::runtime_error("gather tiles: grid size not a multiple of the tile size");
  if (k < 1 || k > 65536) throw std::runtime_error("gather tiles: chunk size out of range");
  const int nt = n / ts, ntile = nt * nt, tt = ts * ts;
  g->ts = ts; g->k = k; g->ntile = ntile;
  // tile -> nodes meeting it, ascending (nodes are visited in order, and a
  // node is appended to a tile at most once)
  std::vector<std::vector<int>> tn(ntile);
  std::vector<int> mark(ntile, -1);
  for (int m = 0; m < M; m++)
    for (int a = 0; a < W; a++) {
      int ty = pmod(t.c0[m] + a, n) / ts;
      for (int b = 0; b < W; b++) {
        int tl = ty * nt + pmod(t.c1[m] + b, n) / ts;
        if (mark[tl] != m) { mark[tl] = m; tn[tl].push_back(m); }
      }
    }
  g->tile_chunk.assign(ntile + 1, 0);
  g->chunk_node.assign(1, 0);
  g->node.clear();
  for (int tl = 0; tl < ntile; tl++) {
    const std::vector<int> &v = tn[tl];
    int nch = ((int)v.size() + k - 1) / k;
    for (int c = 0; c < nch; c++) {
      int a = c * k, b = std::min((int)v.size(), a + k);
      g->node.insert(g->node.end(), v.begin() + a, v.begin() + b);
      g->chunk_node.push_back((int)g->node.size());
    }
    g->tile_chunk[tl + 1] = g->tile_chunk[tl] + nch;
  }
  const int nchunk = g->tile_chunk[ntile];
  g->chunk_tile.assign(nchunk, 0);
  g->chunk_part.assign(nchunk, -1);
  g->comb.clear();
  g->npart = 0;
  for (int tl = 0; tl < ntile; tl++) {
    const int cb = g->tile_chunk[tl], ce = g->tile_chunk[tl + 1];
    for (int c = cb; c < ce; c++) g->chunk_tile[c] = tl;
    if (ce - cb != 1) {
      g->comb.push_back(tl);
      g->comb.push_back(g->npart);
      g->comb.push_back(ce - cb);
      for (int c = cb; c < ce; c++) g->chunk_part[c] = g->npart++;
    }
  }
  g->chunk_cell.assign((size_t)nchunk * (tt + 1), 0);
  g->slot.clear();
  g->weight.clear();
  g->max_chunk_entries = 0;
  // per chunk and cell: the cell's taps (already in node order in the
  // per-cell map) whose node lies in the chunk
  for (int tl = 0; tl < ntile; tl++) {
    const int ty = tl / nt, tx = tl % nt;
    const int cb = g->tile_chunk[tl], ce = g->tile_chunk[tl + 1];
    for (int c = cb; c < ce; c++) {
      const int *nb = g->node.data() + g->chunk_node[c], *ne = g->node.data() + g->chunk_node[c + 1];
      const size_t e_start = g->slot.size();
      int *cc = g->chunk_cell.data() + (size_t)c * (tt + 1);
      for (int cl = 0; cl < tt; cl++) {
        cc[cl] = (int)g->slot.size();
        size_t cell = (size_t)(ty * ts + cl / ts) * n + (size_t)(tx * ts + cl % ts);
        for (int e = t.cell_ptr[cell]; e < t.cell_ptr[cell + 1]; e++) {
          uint32_t u = t.ent[e];
          int m = (int)(u & 0xFFFFFu), t0 = (int)((u >> 20) & 63u), t1 = (int)(u >> 26);
          if (m < *nb || m > *(ne - 1)) continue;
          const int *it = std::lower_bound(nb, ne, m);
          g->slot.push_back((uint16_t)(it - nb));
          g->weight.push_back(t.psi0[(size_t)m * W + t0] * t.psi1[(size_t)m * W + t1]);
        }
      }
      cc[tt] = (int)g->slot.size();
      g->max_chunk_entries = std::max(g->max_chunk_entries, (int)(g->slot.size() - e_start));
    }
  }
  if (g->slot.size() > (size_t)INT32_MAX) throw std::runtime_error("gather tiles: too many entries");
}

int next_pow2(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

int nice_fft_size(int v) {
  for (int s = (v < 2 ? 2 : v + (v & 1));; s += 2) {
    int m = s;
    while (m % 2 == 0) m /= 2;
    while (m % 3 == 0) m /= 3;
    while (m % 5 == 0) m /= 5;
    if (m == 1) return s;
  }
}

Mt19937::Mt19937(unsigned long s) {
  if (s == 0) s = 4357;
  mt[0] = (uint32_t)(s & 0xffffffffUL);
  for (int i = 1; i < 624; i++) mt[i] = (uint32_t)(1812433253UL * (mt[i - 1] ^ (mt[i - 1] >> 30)) + i);
  mti = 624;
}

uint32_t Mt19937::next() {
  const uint32_t UPPER = 0x80000000u, LOWER = 0x7fffffffu;
  if (mti >= 624) {
    int k;
    for (k = 0; k < 624 - 397; k++) {
      uint32_t y = (mt[k] & UPPER) | (mt[k + 1] & LOWER);
      mt[k] = mt[k + 397] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
    }
    for (; k < 623; k++) {
      uint32_t y = (mt[k] & UPPER) | (mt[k + 1] & LOWER);
      mt[k] = mt[k + (397 - 624)] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
    }
    uint32_t y = (mt[623] & UPPER) | (mt[0] & LOWER);
    mt[623] = mt[396] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
    mti = 0;
  }
  uint32_t y = mt[mti++];
  y ^= (y >> 11);
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= (y >> 18);
  return y;
}

}  // namespace nudim
