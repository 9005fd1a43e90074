// fft.hpp -- batched complex FFTs in LDS for gfx950.
//
// Stockham autosort, radices 2, 3, 4, 5 and 8.  A workgroup of 256 threads
// transforms G lines of length n held in LDS as lds[k*G + g] (the G lines
// are G neighbouring planes, which are contiguous in HBM in the
// plane-innermost layout).  Each stage reads its butterfly inputs into
// registers, synchronises, and writes the outputs back in place.
//
// Stage s with radix R and l = product of earlier radices, m = n/(l R):
//   out[(k + l u) m + q] = sum_t w_R^{t u} (w_{lR}^{t k} in[(k R + t) m + q])
// for k < l, q < m; the result is in natural order after the last stage.
#pragma once

#include "common.hpp"

namespace nudim {

constexpr int kFftThreads = 256;

struct FftDesc {
  int n;
  int nf;
  int f[20];
};

FftDesc make_fft_desc(int n);  // throws for lengths with prime factors > 5
// exp(-2 pi i j / n), j < n, in double
void fft_twiddles(int n, std::vector<double> *tw);

template <typename T, int R>
__device__ __forceinline__ void dft_small(cx<T> *v, int sign) {
  if constexpr (R == 2) {
    cx<T> a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
  } else if constexpr (R == 3) {
    const T h = (T)0.5, s3 = (T)0.86602540378443864676;
    cx<T> t1 = v[1] + v[2], t2 = v[1] - v[2];
    cx<T> a = v[0] - h * t1;
    cx<T> b = mul_i(s3 * t2, sign);
    v[0] = v[0] + t1;
    v[1] = a + b;
    v[2] = a - b;
  } else if constexpr (R == 4) {
    cx<T> s02 = v[0] + v[2], d02 = v[0] - v[2];
    cx<T> s13 = v[1] + v[3], d13 = mul_i(v[1] - v[3], sign);
    v[0] = s02 + s13;
    v[1] = d02 + d13;
    v[2] = s02 - s13;
    v[3] = d02 - d13;
  } else if constexpr (R == 5) {
    const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
    const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
    cx<T> b1 = v[1] + v[4], b2 = v[2] + v[3], d1 = v[1] - v[4], d2 = v[2] - v[3];
    cx<T> a0 = v[0];
    cx<T> r1 = a0 + c1 * b1 + c2 * b2;
    cx<T> r2 = a0 + c2 * b1 + c1 * b2;
    cx<T> i1 = mul_i(s1 * d1 + s2 * d2, sign);
    cx<T> i2 = mul_i(s2 * d1 - s1 * d2, sign);
    v[0] = a0 + b1 + b2;
    v[1] = r1 + i1;
    v[4] = r1 - i1;
    v[2] = r2 + i2;
    v[3] = r2 - i2;
  } else if constexpr (R == 8) {
    const T r2 = (T)0.70710678118654752440;
    cx<T> e[4] = {v[0], v[2], v[4], v[6]};
    cx<T> o[4] = {v[1], v[3], v[5], v[7]};
    dft_small<T, 4>(e, sign);
    dft_small<T, 4>(o, sign);
    // w8^1 = r2 (1 + sign i), w8^2 = sign i, w8^3 = r2 (-1 + sign i)
    cx<T> o1 = o[1] * mk(r2, (T)sign * r2);
    cx<T> o2 = mul_i(o[2], sign);
    cx<T> o3 = o[3] * mk(-r2, (T)sign * r2);
    v[0] = e[0] + o[0];
    v[4] = e[0] - o[0];
    v[1] = e[1] + o1;
    v[5] = e[1] - o1;
    v[2] = e[2] + o2;
    v[6] = e[2] - o2;
    v[3] = e[3] + o3;
    v[7] = e[3] - o3;
  }
}

// k = bf / m, q = bf % m for bf < 2^20 by a float reciprocal and one fix-up
__device__ __forceinline__ void divmod_small(int bf, int m, float inv_m, int &k, int &q) {
  k = (int)((float)bf * inv_m);
  q = bf - k * m;
  if (q < 0) {
    k--;
    q += m;
  } else if (q >= m) {
    k++;
    q -= m;
  }
}

// one radix-R stage over G lines (Gs = log2 G) of length n in LDS; the
// twiddle table tw (w_n^j, j < n) is in LDS too
template <typename T, int R, int EPT>
__device__ __forceinline__ void fft_stage(cx<T> *lds, int n, int Gs, int l, const cx<T> *tw, int sign) {
  const int m = n / (l * R);
  const float inv_m = 1.0f / (float)m;
  const int nb = (n / R) << Gs;
  constexpr int BPT = (EPT + R - 1) / R;
  const int G1 = (1 << Gs) - 1;
  cx<T> v[BPT][R];
  const int tid = threadIdx.x;
#pragma unroll
  for (int j = 0; j < BPT; j++) {
    int b = tid + j * kFftThreads;
    if (b < nb) {
      int g = b & G1, bf = b >> Gs;
      int k, q;
      divmod_small(bf, m, inv_m, k, q);
      const int base = ((k * R) * m + q) << Gs;
      const int step = m << Gs;
      const int tstep = k * m;
#pragma unroll
      for (int t = 0; t < R; t++) {
        cx<T> a = lds[base + t * step + g];
        if (t > 0) {
          cx<T> w = tw[t * tstep];
          if (sign > 0) w = conjx(w);
          a = a * w;
        }
        v[j][t] = a;
      }
      dft_small<T, R>(v[j], sign);
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < BPT; j++) {
    int b = tid + j * kFftThreads;
    if (b < nb) {
      int g = b & G1, bf = b >> Gs;
      int k, q;
      divmod_small(bf, m, inv_m, k, q);
      const int base = (k * m + q) << Gs;
      const int step = (l * m) << Gs;
#pragma unroll
      for (int u = 0; u < R; u++) lds[base + u * step + g] = v[j][u];
    }
  }
  __syncthreads();
}

template <typename T, int EPT>
__device__ void fft_lds(cx<T> *lds, const FftDesc &fd, int Gs, const cx<T> *tw, int sign) {
  int l = 1;
  for (int s = 0; s < fd.nf; s++) {
    int r = fd.f[s];
    switch (r) {
      case 8: fft_stage<T, 8, EPT>(lds, fd.n, Gs, l, tw, sign); break;
      case 4: fft_stage<T, 4, EPT>(lds, fd.n, Gs, l, tw, sign); break;
      case 2: fft_stage<T, 2, EPT>(lds, fd.n, Gs, l, tw, sign); break;
      case 3: fft_stage<T, 3, EPT>(lds, fd.n, Gs, l, tw, sign); break;
      default: fft_stage<T, 5, EPT>(lds, fd.n, Gs, l, tw, sign); break;
    }
    l *= r;
  }
}

// copies the n-entry twiddle table into LDS (caller synchronises)
template <typename T>
__device__ __forceinline__ void load_twiddles(cx<T> *stw, const cx<T> *__restrict__ tw, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) stw[i] = tw[i];
}

__device__ __forceinline__ int wrap_pos(int pos, int n) {
  pos += pos < 0 ? n : 0;
  pos -= pos >= n ? n : 0;
  return pos;
}

// Elements per thread of the register stash: n*G <= 256*EPT.
template <typename T>
struct FftEpt { static constexpr int value = sizeof(T) == 4 ? 32 : 16; };

// Addressing of one side of a line transform.  Element k (< cnt) of line
// (o, g) lives at base[o*so + k*sk + chunk*G + g] and sits at FFT position
// (k + shift) mod n, scaled by scale_k[k] * scale_o[o] when given.
struct LineIO {
  long long so, sk;
  int cnt, shift;
  const double *scale_k;
  const double *scale_o;
};

// Passes of the Toeplitz normal operator on the fixed-length kernels:
// row forward (with the CGNR prologue), column (forward, product, inverse),
// row inverse (with the <p, q> partial sums), and a column pass over two
// inputs with two kernels (the difference of two Toeplitz products).
enum { TK_ROWF = 0, TK_COL = 1, TK_ROWI = 2, TK_COL2 = 3 };

// CGNR vectors carried by the row passes, [N][N][P] with planes innermost.
//   row forward: p = z (zz_old null) or p = (zz/zz_old) p + z, after
//                f += alpha p with the old p (alpha non-null); then FFT of p.
//                With z null the pass transforms its input as it is.
//   row inverse: q = the cropped inverse (+ add[] when add is non-null);
//                with p non-null part[row][plane] = sum Re conj(p) q
//   two-input column pass: IFFT(mid FFT(in) - mid2 FFT(in2)); with mid3
//                and qpart also qpart[column][plane] = sum over the column
//                of mid3 |FFT(in)|^2 (the quadratic form of a third
//                circulant: in^H C3 in by Parseval)
template <typename T>
struct ToepCg {
  const cx<T> *z;
  cx<T> *p;
  cx<T> *f;
  const double *zz, *zz_old, *alpha;
  double *part;
  const cx<T> *in2;
  const T *mid2;
  const cx<T> *add;
  const T *mid3;
  double *qpart;
  const cx<T> *pin;  // p_{k-1} when the solve keeps each step's p apart (p_k written to p)
  // CG step with the update in the row inverse pass: the column pass leaves
  // the row sums of <p, T p> (sum over the column of mid |P^|^2) in qpart,
  // and the row inverse pass takes z -= alz[plane] q into zu in place of
  // writing q, with the row sums of |z|^2 in part
  cx<T> *zu;
  const double *alz;
};

// Batched line FFT over a plane-innermost array.  Block b (renumbered so
// that neighbouring plane chunks run on the same XCD and share L2 lines)
// handles line o = b / nchunk and planes [chunk*G, chunk*G+G).  With `mid`
// the block multiplies the transform by mid[o*mid_so + pos] and transforms
// back with the opposite sign (fused Toeplitz column pass).  LDS holds the
// n x G lines followed by the n twiddles.
template <typename T>
__global__ void __launch_bounds__(kFftThreads)
k_fft_lines(const cx<T> *in, cx<T> *out, FftDesc fd, LineIO li, LineIO lo,
            const cx<T> *__restrict__ tw, int sign, int Gs, int nchunk, const cx<T> *__restrict__ mid,
            long long mid_so) {
  extern __shared__ unsigned char smem_raw[];
  cx<T> *lds = reinterpret_cast<cx<T> *>(smem_raw);
  constexpr int EPT = FftEpt<T>::value;
  const int n = fd.n, G = 1 << Gs, G1 = G - 1;
  cx<T> *stw = lds + (n << Gs);
  const int lb = (int)xcd_block(blockIdx.x, gridDim.x);
  const int o = lb / nchunk, chunk = lb - o * nchunk;
  const int tid = threadIdx.x;
  load_twiddles(stw, tw, n);
  // zero the positions no input element lands on
  const int nz0 = (n - li.cnt) << Gs;
  for (int i = tid; i < nz0; i += kFftThreads) {
    int g = i & G1, k = li.cnt + (i >> Gs);
    lds[(wrap_pos(k + li.shift, n) << Gs) + g] = mk((T)0, (T)0);
  }
  const cx<T> *src = in + (long long)o * li.so + (long long)chunk * G;
  const T so_scale = li.scale_o ? (T)li.scale_o[o] : (T)1;
  const int nin = li.cnt << Gs;
#pragma unroll 8
  for (int i = tid; i < nin; i += kFftThreads) {
    int g = i & G1, k = i >> Gs;
    cx<T> v = src[(long long)k * li.sk + g];
    if (li.scale_k) v = ((T)li.scale_k[k] * so_scale) * v;
    else if (li.scale_o) v = so_scale * v;
    lds[(wrap_pos(k + li.shift, n) << Gs) + g] = v;
  }
  __syncthreads();
  fft_lds<T, EPT>(lds, fd, Gs, stw, sign);
  if (mid) {
    const cx<T> *mk_ = mid + (long long)o * mid_so;
    for (int i = tid; i < (n << Gs); i += kFftThreads) lds[i] = lds[i] * mk_[i >> Gs];
    __syncthreads();
    fft_lds<T, EPT>(lds, fd, Gs, stw, -sign);
  }
  cx<T> *dst = out + (long long)o * lo.so + (long long)chunk * G;
  const T oo_scale = lo.scale_o ? (T)lo.scale_o[o] : (T)1;
  const int nout = lo.cnt << Gs;
#pragma unroll 8
  for (int i = tid; i < nout; i += kFftThreads) {
    int g = i & G1, k = i >> Gs;
    cx<T> v = lds[(wrap_pos(k + lo.shift, n) << Gs) + g];
    if (lo.scale_k) v = ((T)lo.scale_k[k] * oo_scale) * v;
    else if (lo.scale_o) v = oo_scale * v;
    dst[(long long)k * lo.sk + g] = v;
  }
}

// Start of line `line` of the real-space side of the z transforms: K volumes
// of a batch, vimg elements apart, each [y][x][z]; line = (y x) K + k reads
// volume k's line (y x), and the Fourier side is [y][x][K][z] (plane k nz + p)
__device__ __forceinline__ long long zimg_base(long long line, int n, int K, long long vimg) {
  return K == 1 ? line * n : (line % K) * vimg + (line / K) * n;
}

// z-axis transforms on volumes stored [y][x][z] (z innermost, lines contiguous).
// S(z) = (z + nz/2) mod nz is the reference's z shift (reconstruction.c:536,
// 550, 568, 599, 618).
//
// forward (reconstruction.c:548-559, 567-568, 589):
//   d[z] = img[S(z)];  D = FFT(d);  fh[p] = D[S(p)];  f0[p] = img[S(p)]
// inverse (reconstruction.c:598-599, 606-618):
//   d[k] = g[S(k)];    e = IFFT(d); img[z] = Re(e[S(z)]) / nz
template <typename T>
__global__ void __launch_bounds__(kFftThreads)
k_zfft_fwd(const T *__restrict__ img, cx<T> *__restrict__ fh, cx<T> *__restrict__ f0, FftDesc fd,
           const cx<T> *__restrict__ tw, int Gs, long long nlines, int K, long long vimg) {
  extern __shared__ unsigned char smem_raw[];
  cx<T> *lds = reinterpret_cast<cx<T> *>(smem_raw);
  constexpr int EPT = FftEpt<T>::value;
  const int nz = fd.n, G = 1 << Gs, h = nz / 2;
  cx<T> *stw = lds + (nz << Gs);
  const long long line0 = (long long)blockIdx.x * G;
  const int tid = threadIdx.x;
  load_twiddles(stw, tw, nz);
#pragma unroll 8
  for (int i = tid; i < nz * G; i += kFftThreads) {
    int g = i / nz, z = i - g * nz;
    long long line = line0 + g;
    T v = (T)0;
    if (line < nlines) {
      int sz = z + h;
      sz -= sz >= nz ? nz : 0;
      v = img[zimg_base(line, nz, K, vimg) + sz];
      if (f0) f0[line * nz + z] = mk(v, (T)0);
    }
    lds[(z << Gs) + g] = mk(v, (T)0);
  }
  __syncthreads();
  fft_lds<T, EPT>(lds, fd, Gs, stw, -1);
#pragma unroll 8
  for (int i = tid; i < nz * G; i += kFftThreads) {
    int g = i / nz, p = i - g * nz;
    long long line = line0 + g;
    if (line < nlines) {
      int sp = p + h;
      sp -= sp >= nz ? nz : 0;
      fh[line * nz + p] = lds[(sp << Gs) + g];
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(kFftThreads)
k_zfft_inv(const cx<T> *__restrict__ gin, T *__restrict__ img, FftDesc fd, const cx<T> *__restrict__ tw,
           int Gs, long long nlines, int K, long long vimg) {
  extern __shared__ unsigned char smem_raw[];
  cx<T> *lds = reinterpret_cast<cx<T> *>(smem_raw);
  constexpr int EPT = FftEpt<T>::value;
  const int nz = fd.n, G = 1 << Gs, h = nz / 2;
  cx<T> *stw = lds + (nz << Gs);
  const long long line0 = (long long)blockIdx.x * G;
  const int tid = threadIdx.x;
  load_twiddles(stw, tw, nz);
#pragma unroll 8
  for (int i = tid; i < nz * G; i += kFftThreads) {
    int g = i / nz, k = i - g * nz;
    long long line = line0 + g;
    cx<T> v = mk((T)0, (T)0);
    if (line < nlines) {
      int sk = k + h;
      sk -= sk >= nz ? nz : 0;
      v = gin[line * nz + sk];
    }
    lds[(k << Gs) + g] = v;
  }
  __syncthreads();
  fft_lds<T, EPT>(lds, fd, Gs, stw, +1);
#pragma unroll 8
  for (int i = tid; i < nz * G; i += kFftThreads) {
    int g = i / nz, z = i - g * nz;
    long long line = line0 + g;
    if (line < nlines) {
      int sz = z + h;
      sz -= sz >= nz ? nz : 0;
      img[zimg_base(line, nz, K, vimg) + z] = lds[(sz << Gs) + g].x / (T)nz;
    }
  }
}

// plain batched DFT of contiguous lines
template <typename T>
__global__ void __launch_bounds__(kFftThreads)
k_fft_contig(cx<T> *__restrict__ data, FftDesc fd, const cx<T> *__restrict__ tw, int Gs, long long nlines,
             int sign) {
  extern __shared__ unsigned char smem_raw[];
  cx<T> *lds = reinterpret_cast<cx<T> *>(smem_raw);
  constexpr int EPT = FftEpt<T>::value;
  const int n = fd.n, G = 1 << Gs;
  cx<T> *stw = lds + (n << Gs);
  const long long line0 = (long long)blockIdx.x * G;
  const int tid = threadIdx.x;
  load_twiddles(stw, tw, n);
  for (int i = tid; i < n * G; i += kFftThreads) {
    int g = i / n, k = i - g * n;
    long long line = line0 + g;
    lds[(k << Gs) + g] = line < nlines ? data[line * n + k] : mk((T)0, (T)0);
  }
  __syncthreads();
  fft_lds<T, EPT>(lds, fd, Gs, stw, sign);
  for (int i = tid; i < n * G; i += kFftThreads) {
    int g = i / n, k = i - g * n;
    long long line = line0 + g;
    if (line < nlines) data[line * n + k] = lds[(k << Gs) + g];
  }
}

}  // namespace nudim
