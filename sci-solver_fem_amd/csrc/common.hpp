// Shared host-side helpers: HIP error propagation, owning device buffers,
// launch geometry.  Errors never terminate the process: they throw FemError,
// which the C ABI layer turns into a status code + message.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace fem {

struct FemError : public std::runtime_error {
  int status;
  FemError(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

inline void hip_check(hipError_t e, const char* what, const char* file, int line) {
  if (e != hipSuccess) {
    char buf[512];
    std::snprintf(buf, sizeof(buf), "HIP error %d (%s) in %s at %s:%d", (int)e,
                  hipGetErrorString(e), what, file, line);
    throw FemError(3, buf);
  }
}
#define HIP_CHECK(x) ::fem::hip_check((x), #x, __FILE__, __LINE__)
#define HIP_LAUNCH_CHECK() ::fem::hip_check(hipGetLastError(), "kernel launch", __FILE__, __LINE__)

[[noreturn]] inline void fail(int status, const std::string& msg) { throw FemError(status, msg); }

// Device allocations made on a thread while a DevPoolScope is alive there
// come from the scope's size classes and go back to them when released, so a
// phase with hundreds of temporaries (the AMG setup) neither waits for the
// device at every hipFree nor calls hipMalloc for every temporary.  All work
// of such a phase runs on one stream, which orders a block's reuse after its
// earlier uses.  When the scope ends its free blocks go, after a device
// synchronization, to a per-device reserve of at most 8 GiB that later scopes
// draw from (hipFree of each cost 0.25-0.65 ms, and large frees slow the
// allocations after them); the reserve is freed by pool_flush_reserve
// (fem_release_cached_memory) and whenever an allocation fails.
// Blocks still held when the scope ends (buffers the phase keeps) are freed
// by hipFree whenever their owner releases them.
void* pool_alloc(size_t bytes);   // null when no scope is active on this thread
bool pool_release(void* p);       // false when p is not a free-able pool block
void pool_flush_reserve();
void* dev_malloc(size_t bytes);   // hipMalloc; on failure the reserve is freed and it tries again

class DevPoolScope {
 public:
  DevPoolScope();
  ~DevPoolScope();
  DevPoolScope(const DevPoolScope&) = delete;
  DevPoolScope& operator=(const DevPoolScope&) = delete;

 private:
  void* prev_;
};

// Host -> device copy of pageable host memory in pieces (destination, source,
// bytes).  The HIP runtime stages pageable memory through the calling
// thread, so one thread bounds the rate (29 GB/s at 690 MB); large copies
// are split over kUploadThreads host threads, each on a stream of its own
// kept per device (51 GB/s with four).  Orders after the work queued on s
// and returns when the data is on the device.
struct HostPiece {
  void* dst;
  const void* src;
  size_t bytes;
};
void upload_pieces(const HostPiece* pieces, int npieces, hipStream_t s);

// Owning device allocation.  Moves only.
template <typename T>
class DevBuf {
 public:
  DevBuf() = default;
  explicit DevBuf(size_t n) { resize(n); }
  ~DevBuf() { release(); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
    return *this;
  }
  void resize(size_t n) {
    if (n == n_) return;
    release();
    if (n) {
      p_ = static_cast<T*>(pool_alloc(n * sizeof(T)));
      if (!p_) p_ = static_cast<T*>(dev_malloc(n * sizeof(T)));
    }
    n_ = n;
  }
  // grow-only variant for work areas
  void reserve(size_t n) { if (n > n_) resize(n); }
  void release() {
    if (p_ && !pool_release(p_)) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T* get() const { return p_; }
  size_t size() const { return n_; }
  size_t bytes() const { return n_ * sizeof(T); }
  void upload(const T* src, size_t n, hipStream_t s) {
    resize(n);
    if (n) HIP_CHECK(hipMemcpyAsync(p_, src, n * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void download(T* dst, size_t n, hipStream_t s) const {
    if (n) HIP_CHECK(hipMemcpyAsync(dst, p_, n * sizeof(T), hipMemcpyDeviceToHost, s));
  }
  std::vector<T> to_host(hipStream_t s) const {
    std::vector<T> v(n_);
    download(v.data(), n_, s);
    HIP_CHECK(hipStreamSynchronize(s));
    return v;
  }

 private:
  T* p_ = nullptr;
  size_t n_ = 0;
};

inline unsigned grid_for(int64_t n, int block) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  return (unsigned)g;
}

}  // namespace fem
