// Shared device-layer utilities: error handling, device buffers, wave
// reductions, kernel-function evaluation. gfx950 (CDNA4): 64-lane wavefronts.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "mxc_gpu.h"

namespace mxg {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);
// device address of the allocation-free dense check's scratch, looked up
// once per device at context set-up (the device must be current)
void dense_scratch_setup(int device);

#define MXG_HIP(x)                                                                              \
  do {                                                                                          \
    hipError_t e_ = (x);                                                                        \
    if (e_ != hipSuccess)                                                                       \
      throw ::mxg::Error(MXG_ERR_HIP, std::string(#x) + " failed: " + hipGetErrorString(e_));   \
  } while (0)

#define MXG_REQUIRE(cond, msg)                                   \
  do {                                                           \
    if (!(cond)) throw ::mxg::Error(MXG_ERR_ARG, (msg));         \
  } while (0)

// Run a C-ABI body, converting exceptions into status codes + last error.
template <class F>
int guarded(F&& f) {
  try {
    f();
    return MXG_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_last_error("host allocation failed");
    return MXG_ERR_NOMEM;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return MXG_ERR_ARG;
  }
}

// Owning device buffer.
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    release();
    if (count == 0) return;
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      throw Error(MXG_ERR_NOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed: " + hipGetErrorString(e));
    }
    n = count;
  }
  void alloc_zero(size_t count, hipStream_t s) {
    alloc(count);
    if (n) MXG_HIP(hipMemsetAsync(p, 0, n * sizeof(T), s));
  }
  void upload(const T* src, size_t count, hipStream_t s) {
    alloc(count);
    if (count) MXG_HIP(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void upload(const std::vector<T>& v, hipStream_t s) { upload(v.data(), v.size(), s); }
  // grow-only scratch
  void reserve(size_t count) {
    if (count > n) alloc(count);
  }
};

// Device scratch handed out in order and taken back all at once by rewind(),
// which first waits for the stream. Its blocks stay allocated for later use,
// so a sequence of requests that repeats (a matrix factorized again, or in a
// lower precision) calls no hipMalloc: allocations of several GB have taken
// up to seconds each on this runtime when large buffers had been freed
// shortly before. Contents are not cleared between uses.
class Arena {
 public:
  Arena() = default;
  Arena(const Arena&) = delete;
  Arena& operator=(const Arena&) = delete;
  ~Arena() {
    for (Block& b : blocks_) (void)hipFree(b.p);
  }
  void rewind(hipStream_t s) {
    MXG_HIP(hipStreamSynchronize(s));
    cur_ = 0;
    off_ = 0;
  }
  void* take(size_t bytes) {
    const size_t need = ((bytes ? bytes : 1) + 255) & ~size_t(255);
    while (cur_ < blocks_.size() && off_ + need > blocks_[cur_].cap) {
      cur_++;
      off_ = 0;
    }
    if (cur_ == blocks_.size()) {
      constexpr size_t G = size_t(2) << 20;
      Block b;
      b.cap = (need + G - 1) / G * G;
      hipError_t e = hipMalloc(reinterpret_cast<void**>(&b.p), b.cap);
      if (e != hipSuccess)
        throw Error(MXG_ERR_NOMEM, "hipMalloc of " + std::to_string(b.cap) + " bytes of scratch failed: " + hipGetErrorString(e));
      blocks_.push_back(b);
      off_ = 0;
    }
    char* p = blocks_[cur_].p + off_;
    off_ += need;
    return p;
  }
  // device bytes held
  size_t held() const {
    size_t t = 0;
    for (const Block& b : blocks_) t += b.cap;
    return t;
  }

 private:
  struct Block {
    char* p = nullptr;
    size_t cap = 0;
  };
  std::vector<Block> blocks_;
  size_t cur_ = 0, off_ = 0;
};

// Descriptor arrays (task lists) of a sequence of launches on one stream. An
// array is copied into pinned host memory and from there, by an asynchronous
// copy on the stream, into device memory; both stay valid until rewind(),
// which first waits for the stream. Launches that read them need no
// synchronization of their own, so the host builds a launch's descriptors
// while the GPU runs the earlier launches. Blocks are kept for later use.
class Stage {
 public:
  Stage() = default;
  Stage(const Stage&) = delete;
  Stage& operator=(const Stage&) = delete;
  ~Stage() {
    for (Block& b : blocks_) {
      if (b.d) (void)hipFree(b.d);
      if (b.h) (void)hipHostFree(b.h);
    }
  }
  void rewind(hipStream_t s) {
    MXG_HIP(hipStreamSynchronize(s));
    cur_ = 0;
    off_ = 0;
  }
  // device copy of v (nullptr for an empty v), valid until rewind()
  template <class X>
  const X* put(const std::vector<X>& v, hipStream_t s) {
    return static_cast<const X*>(put_bytes(v.data(), v.size() * sizeof(X), s));
  }

 private:
  struct Block {
    char* h = nullptr;
    char* d = nullptr;
    size_t cap = 0;
  };
  const void* put_bytes(const void* src, size_t bytes, hipStream_t s) {
    if (bytes == 0) return nullptr;
    const size_t need = (bytes + 255) & ~size_t(255);
    while (cur_ < blocks_.size() && off_ + need > blocks_[cur_].cap) {
      cur_++;
      off_ = 0;
    }
    if (cur_ == blocks_.size()) {
      Block b;
      b.cap = need > (size_t(32) << 20) ? need : (size_t(32) << 20);
      hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&b.h), b.cap, hipHostMallocDefault);
      if (e != hipSuccess) throw Error(MXG_ERR_NOMEM, std::string("pinned staging allocation failed: ") + hipGetErrorString(e));
      e = hipMalloc(reinterpret_cast<void**>(&b.d), b.cap);
      if (e != hipSuccess) {
        (void)hipHostFree(b.h);
        throw Error(MXG_ERR_NOMEM, std::string("staging allocation failed: ") + hipGetErrorString(e));
      }
      blocks_.push_back(b);
      off_ = 0;
    }
    Block& b = blocks_[cur_];
    std::memcpy(b.h + off_, src, bytes);
    MXG_HIP(hipMemcpyAsync(b.d + off_, b.h + off_, bytes, hipMemcpyHostToDevice, s));
    const void* p = b.d + off_;
    off_ += need;
    return p;
  }
  std::vector<Block> blocks_;
  size_t cur_ = 0, off_ = 0;
};

inline unsigned ceil_div(long long a, long long b) { return (unsigned)((a + b - 1) / b); }

// ---- device helpers ----

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ inline float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Kernel function K(d) from the squared distance (kernel.hpp:28-61).
// d = sqrt(d2) and 1/d are correctly rounded IEEE operations; exp is the
// device libm exp (within a few ulp of the host's).
__device__ inline double kernel_eval(int id, double diag, double alpha, double d2) {
  if (id == MXG_KERNEL_GAUSSIAN) return d2 == 0. ? 1.0001 : exp(-alpha * d2);
  if (d2 == 0.) return diag;
  double d = sqrt(d2);
  if (id == MXG_KERNEL_YUKAWA) return exp(-alpha * d) / d;
  return 1. / d;
}

__device__ inline double dist2(const double* a, const double* b) {
  double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  return dx * dx + dy * dy + dz * dz;
}

// Binary search: largest t with prefix[t] <= x (prefix non-decreasing, prefix[0] = 0).
__device__ inline int find_task(const long long* prefix, int ntasks, long long x) {
  int lo = 0, hi = ntasks - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= x) lo = mid; else hi = mid - 1;
  }
  return lo;
}

}  // namespace mxg

// Context object (opaque in the C ABI).
struct mxg_ctx_s {
  int device = 0;
  hipStream_t stream = nullptr;
  // a second stream for work that may overlap the context stream's (the
  // leaf near products of a matvec), forked from and joined back into it
  hipStream_t aux = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  // kernel timing: while enabled, instrumented launches are bracketed by HIP
  // events on the context stream (one start/stop pair per launch)
  bool profiling = false;
  struct Timed {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    long long launches = 0;
    double ms = 0.;
    double work = 0.;  // algorithmic work of the launches, where counted
  };
  Timed timed[MXG_PROF_CLASSES];
  std::vector<hipEvent_t> event_pool;
  // GMRES workspace, grown on demand and kept for later calls on the context
  // (the FP64 Krylov bases of 8 right-hand sides at N = 2^20 and restart 30
  // are 2 GB); calls on one context run one after another
  struct GmresWork {
    mxg::DevBuf<double> V, w, r, part, hd, yd, vcur;
    mxg::DevBuf<float> V32;
    mxg::DevBuf<int> kd;
    // Hessenberg matrices, least-squares scratch, rotations, rotated
    // right-hand sides, thresholds and flags of the right-hand sides
    mxg::DevBuf<double> H, W, cs, sn, gv, tb;
    mxg::DevBuf<int> flags;
    double* hpin = nullptr;  // pinned host copy of one Arnoldi step
    size_t hpin_n = 0;
  } gmres;
  // descriptor staging and per-level scratch of the factorizations on the
  // context stream
  mxg::Stage stage;
  mxg::Arena fscratch;
};

namespace mxg {
// Brackets one launch of kernel class `cls` with events when profiling, on
// the context stream or on stream s.
struct ProfScope {
  mxg_ctx_s* c;
  int cls;
  hipStream_t st;
  hipEvent_t stop = nullptr;
  ProfScope(mxg_ctx_s* ctx, int k, hipStream_t s = nullptr, double work = 0.);
  ~ProfScope() { end(); }
  // records the stop event now (once; the destructor does it otherwise)
  void end();
};
}  // namespace mxg
