// Context, memory helpers and error reporting of the device layer.
// Replaces gpu_handles.cu:5-12 (device selection) and :54-71 (streams and
// handles); there are no BLAS/SPARSE handles: every kernel is our own.
#include "common.hpp"

namespace mxg {
namespace {
thread_local std::string g_last_error;
}
void set_last_error(const std::string& msg) { g_last_error = msg; }

namespace {
hipEvent_t take_event(mxg_ctx_s* c) {
  if (!c->event_pool.empty()) {
    hipEvent_t e = c->event_pool.back();
    c->event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  MXG_HIP(hipEventCreate(&e));
  return e;
}

// CUs kept free of the second stream's work: the short dependent launches of
// the context stream (the matvec's upward pass, one level at a time) start on
// them at once instead of waiting for CUs that the leaf near products hold.
// The runtime deals the mask bits out to the XCDs in turn (bit i to XCD i mod
// 8 on a 256-CU MI355X), so clearing the lowest 32 bits frees 4 CUs on every
// XCD; a mask that leaves some XCD no CU at all is not valid.
constexpr int AUX_RESERVED_CUS = 32;

// The second stream, masked to every CU but the first AUX_RESERVED_CUS, on a
// device with more than 4 AUX_RESERVED_CUS CUs (the whole MI355X); a plain
// stream on smaller partitions or where the runtime refuses the mask.
hipError_t create_aux_stream(int device, hipStream_t* s) {
  int ncu = 0;
  hipError_t e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device);
  if (e == hipSuccess && AUX_RESERVED_CUS > 0 && ncu > 4 * AUX_RESERVED_CUS) {
    std::vector<uint32_t> mask((ncu + 31) / 32, 0u);
    for (int cu = AUX_RESERVED_CUS; cu < ncu; cu++) mask[cu / 32] |= 1u << (cu % 32);
    if (hipExtStreamCreateWithCUMask(s, (uint32_t)mask.size(), mask.data()) == hipSuccess) return hipSuccess;
    (void)hipGetLastError();
  }
  return hipStreamCreateWithFlags(s, hipStreamNonBlocking);
}

void drain(mxg_ctx_s* c) {
  MXG_HIP(hipStreamSynchronize(c->stream));
  MXG_HIP(hipStreamSynchronize(c->aux));
  for (auto& t : c->timed) {
    for (auto& pr : t.pending) {
      float ms = 0.f;
      MXG_HIP(hipEventElapsedTime(&ms, pr.first, pr.second));
      t.ms += ms;
      t.launches++;
      c->event_pool.push_back(pr.first);
      c->event_pool.push_back(pr.second);
    }
    t.pending.clear();
  }
}
}  // namespace

ProfScope::ProfScope(mxg_ctx_s* ctx, int k, hipStream_t s, double work) : c(ctx), cls(k), st(s ? s : ctx->stream) {
  if (!c->profiling || cls < 0 || cls >= MXG_PROF_CLASSES) return;
  c->timed[cls].work += work;
  hipEvent_t start = take_event(c);
  stop = take_event(c);
  MXG_HIP(hipEventRecord(start, st));
  c->timed[cls].pending.emplace_back(start, stop);
}

void ProfScope::end() {
  if (stop) (void)hipEventRecord(stop, st);
  stop = nullptr;
}
}  // namespace mxg

using namespace mxg;

extern "C" {

const char* mxg_last_error(void) { return g_last_error.c_str(); }

int mxg_device_count(int* count) {
  return guarded([&] {
    MXG_REQUIRE(count, "count is null");
    MXG_HIP(hipGetDeviceCount(count));
  });
}

int mxg_init(int device, mxg_ctx_t* ctx) {
  return guarded([&] {
    MXG_REQUIRE(ctx, "ctx is null");
    int n = 0;
    MXG_HIP(hipGetDeviceCount(&n));
    MXG_REQUIRE(n > 0, "no HIP device visible");
    MXG_REQUIRE(device >= 0 && device < n, "device index out of range");
    MXG_HIP(hipSetDevice(device));
    auto* c = new mxg_ctx_s();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = create_aux_stream(device, &c->aux);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->join, hipEventDisableTiming);
    if (e != hipSuccess) {
      if (c->join) (void)hipEventDestroy(c->join);
      if (c->fork) (void)hipEventDestroy(c->fork);
      if (c->aux) (void)hipStreamDestroy(c->aux);
      if (c->stream) (void)hipStreamDestroy(c->stream);
      delete c;
      throw Error(MXG_ERR_HIP, std::string("stream or event creation failed: ") + hipGetErrorString(e));
    }
    try {
      dense_scratch_setup(device);
    } catch (...) {
      (void)hipEventDestroy(c->join);
      (void)hipEventDestroy(c->fork);
      (void)hipStreamDestroy(c->aux);
      (void)hipStreamDestroy(c->stream);
      delete c;
      throw;
    }
    *ctx = c;
  });
}

int mxg_profile_enable(mxg_ctx_t ctx, int enable) {
  return guarded([&] {
    MXG_REQUIRE(ctx, "ctx is null");
    ctx->profiling = enable != 0;
  });
}

int mxg_profile_reset(mxg_ctx_t ctx) {
  return guarded([&] {
    MXG_REQUIRE(ctx, "ctx is null");
    drain(ctx);
    for (auto& t : ctx->timed) {
      t.launches = 0;
      t.ms = 0.;
      t.work = 0.;
    }
  });
}

int mxg_profile_read(mxg_ctx_t ctx, int cls, long long* launches, double* total_ms) {
  return guarded([&] {
    MXG_REQUIRE(ctx && launches && total_ms, "null argument");
    MXG_REQUIRE(cls >= 0 && cls < MXG_PROF_CLASSES, "kernel class out of range");
    drain(ctx);
    *launches = ctx->timed[cls].launches;
    *total_ms = ctx->timed[cls].ms;
  });
}

int mxg_profile_work(mxg_ctx_t ctx, int cls, double* work) {
  return guarded([&] {
    MXG_REQUIRE(ctx && work, "null argument");
    MXG_REQUIRE(cls >= 0 && cls < MXG_PROF_CLASSES, "kernel class out of range");
    *work = ctx->timed[cls].work;
  });
}

int mxg_finalize(mxg_ctx_t ctx) {
  return guarded([&] {
    if (!ctx) return;
    MXG_HIP(hipSetDevice(ctx->device));
    drain(ctx);
    for (hipEvent_t e : ctx->event_pool) (void)hipEventDestroy(e);
    if (ctx->gmres.hpin) MXG_HIP(hipHostFree(ctx->gmres.hpin));
    ctx->gmres.hpin = nullptr;
    MXG_HIP(hipStreamSynchronize(ctx->stream));
    MXG_HIP(hipStreamSynchronize(ctx->aux));
    MXG_HIP(hipEventDestroy(ctx->fork));
    MXG_HIP(hipEventDestroy(ctx->join));
    MXG_HIP(hipStreamDestroy(ctx->aux));
    MXG_HIP(hipStreamDestroy(ctx->stream));
    delete ctx;
  });
}

void* mxg_stream(mxg_ctx_t ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int mxg_mem_info(mxg_ctx_t ctx, size_t* free_bytes, size_t* total_bytes) {
  return guarded([&] {
    MXG_REQUIRE(ctx && free_bytes && total_bytes, "null argument");
    MXG_HIP(hipSetDevice(ctx->device));
    MXG_HIP(hipMemGetInfo(free_bytes, total_bytes));
  });
}

int mxg_synchronize(mxg_ctx_t ctx) {
  return guarded([&] {
    MXG_REQUIRE(ctx, "ctx is null");
    MXG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int mxg_malloc(mxg_ctx_t ctx, size_t bytes, void** d_ptr) {
  return guarded([&] {
    MXG_REQUIRE(ctx && d_ptr, "null argument");
    MXG_HIP(hipSetDevice(ctx->device));
    *d_ptr = nullptr;
    if (bytes == 0) return;
    hipError_t e = hipMalloc(d_ptr, bytes);
    if (e != hipSuccess) throw Error(MXG_ERR_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
  });
}

int mxg_free(mxg_ctx_t ctx, void* d_ptr) {
  return guarded([&] {
    MXG_REQUIRE(ctx, "ctx is null");
    if (d_ptr) MXG_HIP(hipFree(d_ptr));
  });
}

int mxg_memcpy_h2d(mxg_ctx_t ctx, void* d_dst, const void* src, size_t bytes) {
  return guarded([&] {
    MXG_REQUIRE(ctx, "ctx is null");
    if (bytes == 0) return;
    MXG_REQUIRE(d_dst && src, "null pointer");
    MXG_HIP(hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    MXG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int mxg_memcpy_d2h(mxg_ctx_t ctx, void* dst, const void* d_src, size_t bytes) {
  return guarded([&] {
    MXG_REQUIRE(ctx, "ctx is null");
    if (bytes == 0) return;
    MXG_REQUIRE(dst && d_src, "null pointer");
    MXG_HIP(hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    MXG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int mxg_memcpy_d2d(mxg_ctx_t ctx, void* d_dst, const void* d_src, size_t bytes) {
  return guarded([&] {
    MXG_REQUIRE(ctx, "ctx is null");
    if (bytes == 0) return;
    MXG_REQUIRE(d_dst && d_src, "null pointer");
    MXG_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    MXG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int mxg_memset(mxg_ctx_t ctx, void* d_dst, int value, size_t bytes) {
  return guarded([&] {
    MXG_REQUIRE(ctx, "ctx is null");
    if (bytes == 0) return;
    MXG_REQUIRE(d_dst, "null pointer");
    MXG_HIP(hipMemsetAsync(d_dst, value, bytes, ctx->stream));
    MXG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

}  // extern "C"
