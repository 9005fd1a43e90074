// geom.hpp -- host-side geometry of one reconstruction job: node tables,
// density-compensation weights, sample classes, Kaiser-Bessel window tables
// and the cell -> node inverse map of the gather adjoint.  Runs once per job
// at set-up (the reference does the same work on the host at
// src/reconstruction.c:304-426 and inside NFFT3's nfft_init_guru).
#pragma once

#include <cstdint>
#include <vector>

namespace nudim {

// grid.h:11: the reference's value of pi for node angles and weights
constexpr double kRefPi = 3.1415926;

// grid.c:78-97: rays at amin + s*ainterval (accumulated), radius -0.5 + i/np
// (accumulated); x[2m] = r sin(gamma/180*PI) (dimension 0, y),
// x[2m+1] = r cos(gamma/180*PI) (dimension 1, x).  Returns the node count.
int polar_nodes(double amin, double amax, double ainterval, int np, std::vector<double> *x);
// grid.c:100-133
void uniform_nodes(int nx, int nz, std::vector<double> *x);
// reconstruction.c:375-379
void pad_range(double gam_min, double interval, double *pad_min, double *pad_max);
// reconstruction.c:415-426
void density_weights(int np, int nang, double padangle, std::vector<double> *w);
// reconstruction.c:437-484: kind 0 unknown, 1 known (order = tilt file), 2 held out
void sample_classes(int np, int nang, double pad_min, double padangle, double gam_min,
                    double gam_max, double interval, std::vector<int> *kind,
                    std::vector<int> *order);

// Kaiser-Bessel window psi(d) = I0(beta sqrt(1-(2d/W)^2))/I0(beta), |d| <= W/2,
// beta from Beatty, Nishimura and Pauly (IEEE TMI 2005).
double kb_beta(int W, double sigma);
double bessel_i0(double x);
double kb_psi(double d, int W, double beta);
double kb_psihat(double nu, int W, double beta);

// Window tables of a square N x N NFFT on an n x n grid with W taps per
// dimension, for M nodes x[M][2].
struct NfftTables {
  int N = 0, n = 0, W = 0, M = 0;
  std::vector<int> c0, c1;          // first tap cell per node and dimension (unwrapped)
  std::vector<double> psi0, psi1;   // [M][W]
  std::vector<double> dec;          // [N]: 1/psihat((k-N/2)/n)
  // gather inverse map: for each cell, the (node, t0, t1) taps in node order
  std::vector<int> cell_ptr;        // [n*n+1]
  std::vector<uint32_t> ent;        // node | t0 << 20 | t1 << 26
};
void build_nfft_tables(int N, int n, int W, const double *x, int M, bool with_gather,
                       NfftTables *t);

// Patch lists of the spreading kernels.  The n x n grid is cut into patches
// of pr x pc cells (n % pr == n % pc == 0), patch (a, b) = a * (n/pc) + b,
// each held in registers by one wavefront.  A patch's entries are the nodes
// whose W x W window meets it, in increasing node order, each with the
// offsets k0, k1 of the patch's first row and column in the node's
// zero-padded window tables: row i of the patch takes psi0p[node][k0 + i],
// where psi0p[node][k] = psi0[node][k - (pr - 1)] inside the window and 0
// outside (length W + 2 (pr - 1)); columns likewise with pc.
struct PatchLists {
  int pr = 0, pc = 0, npr = 0, npc = 0, l0 = 0, l1 = 0, nonempty = 0;
  std::vector<int> ptr;             // [npatch + 1]: entry range of each patch
  std::vector<uint32_t> ent;        // node | k0 << 20 | k1 << 26
  std::vector<int> order;           // patches by entry count, largest first (ties: patch id)
  std::vector<double> psi0p, psi1p; // [M][l0], [M][l1]
  std::vector<int> node_ptr;        // [M + 1]
  std::vector<int> node_ent;        // each node's entry ids, in increasing patch id
};
// false (and pl untouched) when the sizes do not allow patches: n not a
// multiple of pr and pc, n < W + max(pr, pc), offsets past 6 bits, M >= 2^20
bool build_patch_lists(const NfftTables &t, int pr, int pc, PatchLists *pl);

// smallest even 2^a 3^b 5^c >= v
int nice_fft_size(int v);
int next_pow2(int v);

// MT19937 (Matsumoto and Nishimura 1998) with GSL's seeding; uniform() is
// gsl_rng_uniform for gsl_rng_mt19937 (the default generator).
struct Mt19937 {
  uint32_t mt[624];
  int mti;
  explicit Mt19937(unsigned long s);
  uint32_t next();
  double uniform() { return next() / 4294967296.0; }
};

}  // namespace nudim
