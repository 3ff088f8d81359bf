// core.cpp -- device context and caching allocator.
#include "core.hpp"
#include <chrono>
#include <cstring>
#include "eigen.hpp"
#include <rocsolver/rocsolver.h>
#include "kernels.hpp"

namespace sadmrg {

namespace {
thread_local std::string t_last_error;
}

void set_last_error(const std::string& msg) { t_last_error = msg; }
const char* last_error() { return t_last_error.c_str(); }

void fail(int code, const std::string& msg) { throw Error(code, msg); }

DevBuf::DevBuf(Context* c, size_t bytes) : ctx_(c), bytes_(bytes) {
  if (bytes > 0) ptr_ = c->raw_alloc(bytes, &cap_);
}

DevBuf::~DevBuf() { release(); }

void DevBuf::release() {
  if (ptr_ && ctx_) ctx_->raw_free(ptr_, cap_);
  ptr_ = nullptr;
  bytes_ = cap_ = 0;
}

Context::Context(int device) : device_(device) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= device)
    fail(kErrNoDevice, "no HIP device " + std::to_string(device) + " (device count " + std::to_string(n) + ")");
  SADMRG_HIP(hipSetDevice(device));
  // default priority stream (non-default priorities are avoided on purpose)
  SADMRG_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  partial_ = DevBuf(this, sizeof(double) * 2 * kPartialGrid * 64);
  dscal_ = DevBuf(this, sizeof(double) * kMaxDscal);
  noalloc_scratch_setup(device);
}

Context::~Context() {
  if (stream_) hipStreamSynchronize(stream_);
  for (auto& e : staged_) hipEventDestroy(e.ev);
  for (auto e : spare_events_) hipEventDestroy(e);
  if (staging_) hipHostFree(staging_);
  if (solver_) rocblas_destroy_handle(static_cast<rocblas_handle>(solver_));
  partial_.release();
  dscal_.release();
  for (auto& kv : free_) hipFree(kv.second);
  free_.clear();
  if (stream_) hipStreamDestroy(stream_);
}

namespace {
double wall_now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace

void Context::sync() {
  const double t0 = wall_now();
  SADMRG_HIP(hipStreamSynchronize(stream_));
  ++n_sync_;
  t_sync_ += wall_now() - t0;
}

void Context::upload(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return;
  if (bytes > kStagingBytes / 4) {
    SADMRG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream_));
    sync();
    return;
  }
  if (!staging_) SADMRG_HIP(hipHostMalloc(reinterpret_cast<void**>(&staging_), kStagingBytes, hipHostMallocDefault));
  const size_t need = (bytes + 255) / 256 * 256;
  if (staging_head_ + need > kStagingBytes) staging_head_ = 0;
  const size_t b = staging_head_, e = b + need;
  // Wait for the newest in-flight copy that still reads [b, e): copies on the
  // one stream complete in order, so every older one has completed with it.
  int last = -1;
  for (size_t i = 0; i < staged_.size(); ++i)
    if (staged_[i].begin < e && b < staged_[i].end) last = static_cast<int>(i);
  if (last >= 0) SADMRG_HIP(hipEventSynchronize(staged_[last].ev));
  for (int i = 0; i <= last; ++i) {
    spare_events_.push_back(staged_.front().ev);
    staged_.pop_front();
  }
  // drop completed entries from the front so the list stays short
  while (staged_.size() > 64 && hipEventQuery(staged_.front().ev) == hipSuccess) {
    spare_events_.push_back(staged_.front().ev);
    staged_.pop_front();
  }
  std::memcpy(staging_ + b, src, bytes);
  SADMRG_HIP(hipMemcpyAsync(dst, staging_ + b, bytes, hipMemcpyHostToDevice, stream_));
  hipEvent_t ev;
  if (spare_events_.empty()) {
    SADMRG_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  } else {
    ev = spare_events_.back();
    spare_events_.pop_back();
  }
  SADMRG_HIP(hipEventRecord(ev, stream_));
  staged_.push_back(Staged{b, e, ev});
  staging_head_ = e;
}

static size_t round_cap(size_t bytes) {
  const size_t a = 256;
  size_t c = (bytes + a - 1) / a * a;
  if (c >= (size_t(1) << 20)) {  // above 1 MiB: round to 1/8 power-of-two steps
    size_t p = 1;
    while (p < c) p <<= 1;
    const size_t step = p / 8;
    c = (c + step - 1) / step * step;
  }
  return c;
}

void* Context::raw_alloc(size_t bytes, size_t* cap) {
  const size_t c = round_cap(bytes);
  auto it = free_.lower_bound(c);
  if (it != free_.end() && it->first <= c + c / 4) {
    void* p = it->second;
    *cap = it->first;
    cached_ -= it->first;
    in_use_ += it->first;
    free_.erase(it);
    return p;
  }
  void* p = nullptr;
  const double t0 = wall_now();
  hipError_t e = hipMalloc(&p, c);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    trim();
    SADMRG_HIP(hipMalloc(&p, c));
  }
  ++n_malloc_;
  t_malloc_ += wall_now() - t0;
  *cap = c;
  in_use_ += c;
  return p;
}

void Context::raw_free(void* p, size_t cap) {
  in_use_ -= cap;
  cached_ += cap;
  free_.emplace(cap, p);
}

void Context::trim() {
  SADMRG_HIP(hipStreamSynchronize(stream_));
  for (auto& kv : free_) hipFree(kv.second);
  free_.clear();
  cached_ = 0;
}

void* Context::solver() {
  if (!solver_) {
    rocblas_handle h;
    if (rocblas_create_handle(&h) != rocblas_status_success) fail(kErrNumerics, "rocblas_create_handle failed");
    rocblas_set_stream(h, stream_);
    solver_ = h;
  }
  return solver_;
}

}  // namespace sadmrg
