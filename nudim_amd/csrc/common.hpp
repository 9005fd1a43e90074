// common.hpp -- shared types and helpers of libnudim_hip (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace nudim {

// Interleaved complex value; matches fftw_complex for T = double.
template <typename T>
struct cx {
  T x, y;
};

template <typename T>
__host__ __device__ __forceinline__ cx<T> mk(T a, T b) {
  cx<T> r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename T>
__host__ __device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename T>
__host__ __device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return mk(a.x - b.x, a.y - b.y); }
template <typename T>
__host__ __device__ __forceinline__ cx<T> operator*(cx<T> a, cx<T> b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T>
__host__ __device__ __forceinline__ cx<T> operator*(T s, cx<T> a) { return mk(s * a.x, s * a.y); }
template <typename T>
__host__ __device__ __forceinline__ cx<T> conjx(cx<T> a) { return mk(a.x, -a.y); }
// multiply by -i (sign < 0) or +i (sign > 0)
template <typename T>
__host__ __device__ __forceinline__ cx<T> mul_i(cx<T> a, int sign) {
  return sign < 0 ? mk(a.y, -a.x) : mk(-a.y, a.x);
}

// the calling thread's message for nudim_last_error (api.hip)
void set_last_error(const std::string &msg);

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define NUDIM_HIP(call)                                                                  \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      throw ::nudim::Error(-3, std::string(#call) + ": " + hipGetErrorString(e_) + " at " + \
                                   __FILE__ + ":" + std::to_string(__LINE__));             \
  } while (0)

#define NUDIM_CHECK(cond, msg)                                   \
  do {                                                           \
    if (!(cond)) throw ::nudim::Error(-1, std::string(msg));     \
  } while (0)

// Owned device allocation.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  explicit DevBuf(size_t b) { alloc(b); }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t b) {
    release();
    if (b == 0) return;
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) {
      p = nullptr;
      throw Error(-2, "hipMalloc(" + std::to_string(b) + " bytes) failed: " + hipGetErrorString(e));
    }
    bytes = b;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename U>
  U *as() const { return static_cast<U *>(p); }
};

inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// Whether a batch of nvol volumes of nz planes runs each volume with the same
// arithmetic as its own job.  The per-plane reductions group up to 64 planes
// (the largest power of two dividing the plane count, at most 64) and the
// NFFT interpolation takes its patch form (spread.hpp) at plane counts that
// are multiples of 64; nvol nz gives the same choices as nz when nvol is odd
// or nz is a multiple of 64.  (Other batches agree with their own jobs to
// rounding.)
inline bool batch_bit_exact(int nz, int nvol) { return nvol % 2 == 1 || nz % 64 == 0; }

// Kernel classes that can be timed with HIP events on the launching stream.
enum KernelClass {
  KC_NONE = 0,
  KC_TOEPLITZ_COL = 1,  // fused column FFT x kernel x inverse FFT
  KC_INTERP = 2,        // NFFT window interpolation
  KC_GATHER = 3,        // NFFT gather adjoint
  KC_LINE_FFT = 4,      // other line FFT passes
  KC_TV = 5,            // TV sweep kernels
  KC_ZFFT = 6,          // z-axis transforms
  KC_TOEPLITZ_ROWF = 7, // Toeplitz row forward pass with the CG prologue
  KC_TOEPLITZ_ROWI = 8, // Toeplitz row inverse pass with the <p, q> row sums
  KC_TOEPLITZ_COL2 = 9, // two-input column pass of the projected CG start
  KC_CG_UPDATE = 10,    // CG vector update z -= alpha q with |z|^2
  KC_DIFFUSION = 11,    // anisotropic diffusion (all sweeps of one call)
};

// Records an event pair around every launch of one kernel class and
// accumulates the algorithmic bytes the caller attributes to each launch.
struct KernelProfiler {
  int cls = KC_NONE;
  std::vector<hipEvent_t> ev;
  size_t used = 0;
  double bytes = 0, ms = 0;
  long launches = 0;
  bool on(int c) const { return cls != KC_NONE && c == cls; }
  void before(hipStream_t st) {
    if (used + 2 > ev.size()) {
      size_t old = ev.size();
      ev.resize(old + 256);
      for (size_t i = old; i < ev.size(); i++) (void)hipEventCreate(&ev[i]);
    }
    (void)hipEventRecord(ev[used], st);
  }
  void after(hipStream_t st, double b) {
    (void)hipEventRecord(ev[used + 1], st);
    used += 2;
    bytes += b;
    launches++;
  }
  // call after the stream has been synchronised
  void collect() {
    for (size_t i = 0; i + 1 < used; i += 2) {
      float t = 0;
      if (hipEventElapsedTime(&t, ev[i], ev[i + 1]) == hipSuccess) ms += t;
    }
    used = 0;
  }
  void reset(int c) {
    cls = c;
    bytes = ms = 0;
    launches = 0;
    used = 0;
  }
  ~KernelProfiler() {
    for (auto e : ev) (void)hipEventDestroy(e);
  }
};

// profiler of the job driven by the calling thread (null when off)
extern thread_local KernelProfiler *g_prof;

struct ProfScope {
  KernelProfiler *p;
  hipStream_t st;
  double b;
  ProfScope(int c, hipStream_t s, double bytes) : p(g_prof && g_prof->on(c) ? g_prof : nullptr), st(s), b(bytes) {
    if (p) p->before(st);
  }
  ~ProfScope() {
    if (p) p->after(st, b);
  }
};

// Blocks of one launch are dealt round-robin to the 8 XCDs (each with its own
// L2).  Renumber so that each XCD works on one contiguous range of logical
// blocks; neighbouring blocks then share their L2.
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb) {
  const unsigned nx = 8;
  unsigned per = nb / nx, rem = nb % nx, x = b % nx, i = b / nx;
  // XCD x owns per (+1 for x < rem) logical blocks
  unsigned start = x * per + (x < rem ? x : rem);
  unsigned owned = per + (x < rem ? 1 : 0);
  return i < owned ? start + i : b;
}

}  // namespace nudim
