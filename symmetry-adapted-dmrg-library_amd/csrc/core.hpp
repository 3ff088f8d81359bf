// core.hpp -- device context, caching allocator, error handling, scalar traits.
#pragma once
#include <hip/hip_runtime_api.h>
#include <complex>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace sadmrg {

using dcmplex = std::complex<double>;

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

enum Status : int {
  kOk = 0,
  kErrInvalid = 1,     // bad arguments / inconsistent bonds (the reference exits here)
  kErrHip = 2,         // HIP runtime failure
  kErrUnsupported = 3, // a case the reference supports that this engine rejects
  kErrNumerics = 4,    // SVD / eigensolver failure
  kErrNoDevice = 5,
};

[[noreturn]] void fail(int code, const std::string& msg);
inline void require(bool ok, const std::string& msg) {
  if (!ok) fail(kErrInvalid, msg);
}

#define SADMRG_HIP(x)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess)                                                                 \
      ::sadmrg::fail(::sadmrg::kErrHip, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class S> struct Traits;
template <> struct Traits<double> {
  static constexpr bool is_complex = false;
  static constexpr int width = 1;  // doubles per element
};
template <> struct Traits<dcmplex> {
  static constexpr bool is_complex = true;
  static constexpr int width = 2;
};

class Context;

// Owning device allocation returned to the context cache on destruction.
class DevBuf {
 public:
  DevBuf() = default;
  DevBuf(Context* c, size_t bytes);
  ~DevBuf();
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept { swap(o); }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); swap(o); }
    return *this;
  }
  void release();
  void* get() const { return ptr_; }
  template <class T> T* as() const { return static_cast<T*>(ptr_); }
  size_t bytes() const { return bytes_; }
  Context* ctx() const { return ctx_; }

 private:
  void swap(DevBuf& o) noexcept {
    std::swap(ctx_, o.ctx_);
    std::swap(ptr_, o.ptr_);
    std::swap(bytes_, o.bytes_);
    std::swap(cap_, o.cap_);
  }
  Context* ctx_ = nullptr;
  void* ptr_ = nullptr;
  size_t bytes_ = 0;
  size_t cap_ = 0;
};

// One device, one in-order stream.  All work is enqueued on `stream`, so a
// cached allocation can be recycled as soon as the host releases it: later
// users are ordered after earlier ones by the stream.
class Context {
 public:
  explicit Context(int device);
  ~Context();
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;

  int device() const { return device_; }
  hipStream_t stream() const { return stream_; }
  void sync();

  // Host-to-device copy on the stream through pinned staging memory: the
  // source is copied into a staging ring before the call returns, so it may
  // be reused or freed at once, and the host does not wait for the device
  // (a copy from pageable memory would).  The ring (kStagingBytes) is
  // recycled in stream order: a region is reused once the copy that read it
  // has completed.  Larger copies go straight from the source and wait.
  void upload(void* dst, const void* src, size_t bytes);
  static constexpr size_t kStagingBytes = size_t(64) << 20;

  void* raw_alloc(size_t bytes, size_t* cap);
  void raw_free(void* p, size_t cap);
  void trim();

  // scratch for reductions (2 * kPartialGrid * 64 doubles) and device scalars
  double* partial() const { return partial_.as<double>(); }
  double* dscal() const { return dscal_.as<double>(); }
  static constexpr int kMaxDscal = 4096;

  // rocSOLVER handle (created lazily)
  void* solver();

  size_t bytes_in_use() const { return in_use_; }
  size_t bytes_cached() const { return cached_; }
  // host-side counters: device allocations that missed the cache (hipMalloc)
  // and stream synchronisations, with the wall seconds spent in each
  long long mallocs() const { return n_malloc_; }
  double malloc_seconds() const { return t_malloc_; }
  long long syncs() const { return n_sync_; }
  double sync_seconds() const { return t_sync_; }

 private:
  int device_;
  hipStream_t stream_ = nullptr;
  std::multimap<size_t, void*> free_;
  size_t in_use_ = 0, cached_ = 0;
  long long n_malloc_ = 0, n_sync_ = 0;
  double t_malloc_ = 0, t_sync_ = 0;
  DevBuf partial_, dscal_;
  void* solver_ = nullptr;
  // staging ring of upload(): in-flight regions [begin, end) with the event
  // recorded after the copy that reads them, oldest first
  struct Staged {
    size_t begin, end;
    hipEvent_t ev;
  };
  char* staging_ = nullptr;
  size_t staging_head_ = 0;
  std::deque<Staged> staged_;
  std::vector<hipEvent_t> spare_events_;
};

void set_last_error(const std::string& msg);
const char* last_error();

}  // namespace sadmrg
