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

bool build_patch_lists(const NfftTables &t, int pr, int pc, PatchLists *pl) {
  const int n = t.n, W = t.W, M = t.M;
  if (pr < 1 || pc < 1 || n % pr != 0 || n % pc != 0 || n < W + std::max(pr, pc)) return false;
  if (W + std::max(pr, pc) - 2 >= 64 || M >= (1 << 20)) return false;
  const int npr = n / pr, npc = n / pc, l0 = W + 2 * (pr - 1), l1 = W + 2 * (pc - 1);
  const size_t npatch = (size_t)npr * npc;
  // patch rows (columns) a node's window meets, with the window tap of the
  // patch's first row (column): d in [-(p-1), W-1]
  auto bands = [n, W](int c0, int p, std::vector<std::pair<int, int>> *out) {
    out->clear();
    for (int k = 0; k < W; k++) {
      const int band = pmod(c0 + k, n) / p;
      bool seen = false;
      for (auto &b : *out) seen = seen || b.first == band;
      if (seen) continue;
      int d = pmod(band * p - c0, n);
      if (d > W - 1) d -= n;
      out->push_back({band, d});
    }
  };
  std::vector<int> cnt(npatch + 1, 0);
  std::vector<std::pair<int, int>> rb, cb;
  for (int m = 0; m < M; m++) {
    bands(t.c0[m], pr, &rb);
    bands(t.c1[m], pc, &cb);
    for (auto &r : rb)
      for (auto &c : cb) cnt[(size_t)r.first * npc + c.first + 1]++;
  }
  for (size_t q = 0; q < npatch; q++) cnt[q + 1] += cnt[q];
  if (cnt[npatch] > INT32_MAX / 2) return false;
  pl->pr = pr; pl->pc = pc; pl->npr = npr; pl->npc = npc; pl->l0 = l0; pl->l1 = l1;
  pl->ptr.assign(cnt.begin(), cnt.end());
  pl->ent.assign((size_t)cnt[npatch], 0);
  std::vector<int> fill(cnt.begin(), cnt.end() - 1);
  std::vector<int> ecount(M + 1, 0);
  for (int m = 0; m < M; m++) {
    bands(t.c0[m], pr, &rb);
    bands(t.c1[m], pc, &cb);
    for (auto &r : rb)
      for (auto &c : cb) {
        const int k0 = r.second + pr - 1, k1 = c.second + pc - 1;
        pl->ent[fill[(size_t)r.first * npc + c.first]++] = (uint32_t)m | ((uint32_t)k0 << 20) | ((uint32_t)k1 << 26);
      }
    ecount[m + 1] = (int)(rb.size() * cb.size());
  }
  // each node's entries in increasing patch id (patches are visited in order)
  for (int m = 0; m < M; m++) ecount[m + 1] += ecount[m];
  pl->node_ptr = ecount;
  pl->node_ent.assign((size_t)cnt[npatch], 0);
  std::vector<int> nfill(ecount.begin(), ecount.end() - 1);
  for (size_t q = 0; q < npatch; q++)
    for (int e = pl->ptr[q]; e < pl->ptr[q + 1]; e++) pl->node_ent[nfill[pl->ent[e] & 0xFFFFFu]++] = e;
  pl->order.resize(npatch);
  for (size_t q = 0; q < npatch; q++) pl->order[q] = (int)q;
  std::stable_sort(pl->order.begin(), pl->order.end(), [&](int a, int b) {
    return pl->ptr[a + 1] - pl->ptr[a] > pl->ptr[b + 1] - pl->ptr[b];
  });
  pl->nonempty = 0;
  for (size_t q = 0; q < npatch; q++) pl->nonempty += pl->ptr[q + 1] > pl->ptr[q];
  pl->psi0p.assign((size_t)M * l0, 0.0);
  pl->psi1p.assign((size_t)M * l1, 0.0);
  for (int m = 0; m < M; m++)
    for (int k = 0; k < W; k++) {
      pl->psi0p[(size_t)m * l0 + k + pr - 1] = t.psi0[(size_t)m * W + k];
      pl->psi1p[(size_t)m * l1 + k + pc - 1] = t.psi1[(size_t)m * W + k];
    }
  return true;
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
