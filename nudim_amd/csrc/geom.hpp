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

// Tiled form of the gather inverse map.  The n x n grid is cut into tiles of
// ts x ts cells (n % ts == 0), tile (ty, tx) = ty * (n/ts) + tx.  The nodes
// whose W x W footprint meets a tile, in increasing node order, are cut into
// chunks of at most k nodes; a chunk's node values are what one block stages
// in LDS.  Within a chunk the entries are grouped by cell (row-major within
// the tile) and, per cell, kept in node order.  A tile with one chunk is
// summed straight into the grid; the chunks of a tile with several write
// partial sums (slot chunk_part[c]) that a second pass adds in chunk order,
// and that pass also clears the tiles no node meets.
struct GatherTiles {
  int ts = 0, k = 0, ntile = 0, max_chunk_entries = 0, npart = 0;
  std::vector<int> tile_chunk;      // [ntile+1]: chunk ids of each tile
  std::vector<int> chunk_node;      // [nchunk+1]: range of each chunk in node
  std::vector<int> node;            // node ids
  std::vector<int> chunk_cell;      // [nchunk][ts*ts+1]: absolute entry ranges per cell
  std::vector<uint16_t> slot;       // entry -> node slot within its chunk
  std::vector<double> weight;       // entry -> psi0[node][t0] * psi1[node][t1]
  std::vector<int> chunk_tile;      // [nchunk]
  std::vector<int> chunk_part;      // [nchunk]: partial-sum slot, -1 for a tile's only chunk
  std::vector<int> comb;            // tiles with != 1 chunk: (tile, first slot, count) triples
};
void build_gather_tiles(const NfftTables &t, int ts, int k, GatherTiles *g);

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
