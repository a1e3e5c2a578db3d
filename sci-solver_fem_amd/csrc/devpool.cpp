// Setup-phase device allocation pool (see DevPoolScope in common.hpp) and
// the threaded host -> device upload.

#include <algorithm>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "../../include/fem_mi355x.h"
#include "common.hpp"

namespace fem {

namespace {

struct Pool {
  std::unordered_map<void*, size_t> owned;  // blocks this pool allocated -> size class
  std::multimap<size_t, void*> free_;       // released blocks by size class
  int device = 0;
};

thread_local Pool* g_pool = nullptr;

// size classes: 256 B, then 2^k (1 + j / 8): at most 12.5 % above the request
size_t size_class(size_t bytes) {
  if (bytes <= 256) return 256;
  size_t p = 256;
  while (p * 2 <= bytes) p *= 2;
  const size_t step = p / 8;
  return (bytes + step - 1) / step * step;
}

// free blocks of ended scopes, per device, by size class
constexpr size_t kReserveCap = (size_t)8 << 30;
struct Reserve {
  std::multimap<size_t, void*> blocks;
  size_t bytes = 0;
};
std::mutex g_res_mu;
std::unordered_map<int, Reserve> g_res;

}  // namespace

void pool_flush_reserve() {
  std::lock_guard<std::mutex> lk(g_res_mu);
  for (auto& dr : g_res) {
    for (auto& kv : dr.second.blocks) (void)hipFree(kv.second);
    dr.second.blocks.clear();
    dr.second.bytes = 0;
  }
}

void* dev_malloc(size_t bytes) {
  void* p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    pool_flush_reserve();
    HIP_CHECK(hipMalloc(&p, bytes));
  }
  return p;
}

void* pool_alloc(size_t bytes) {
  Pool* pl = g_pool;
  if (!pl || bytes == 0) return nullptr;
  const size_t c = size_class(bytes);
  auto it = pl->free_.find(c);
  if (it != pl->free_.end()) {
    void* p = it->second;
    pl->free_.erase(it);
    return p;
  }
  void* p = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_res_mu);
    auto rt = g_res.find(pl->device);
    if (rt != g_res.end()) {
      auto jt = rt->second.blocks.find(c);
      if (jt != rt->second.blocks.end()) {
        p = jt->second;
        rt->second.blocks.erase(jt);
        rt->second.bytes -= c;
      }
    }
  }
  if (!p) p = dev_malloc(c);
  pl->owned.emplace(p, c);
  return p;
}

bool pool_release(void* p) {
  Pool* pl = g_pool;
  if (!pl) return false;
  auto it = pl->owned.find(p);
  if (it == pl->owned.end()) return false;
  pl->free_.emplace(it->second, p);
  return true;
}

DevPoolScope::DevPoolScope() : prev_(g_pool) {
  Pool* pl = new Pool();
  (void)hipGetDevice(&pl->device);
  g_pool = pl;
}

DevPoolScope::~DevPoolScope() {
  Pool* pl = g_pool;
  if (!pl->free_.empty()) {
    // queued work may still use the blocks: only an idle device hands them on
    const bool idle = hipDeviceSynchronize() == hipSuccess;
    std::lock_guard<std::mutex> lk(g_res_mu);
    Reserve& r = g_res[pl->device];
    for (auto& kv : pl->free_) {
      if (idle && r.bytes + kv.first <= kReserveCap) {
        r.blocks.emplace(kv.first, kv.second);
        r.bytes += kv.first;
      } else {
        (void)hipFree(kv.second);
      }
    }
  }
  delete pl;
  g_pool = static_cast<Pool*>(prev_);
}

namespace {

constexpr int kUploadThreads = 4;
constexpr size_t kUploadSplit = (size_t)8 << 20;  // smaller copies stay one async copy on s
constexpr int kMaxDevices = 64;
std::mutex g_up_mu;
hipStream_t g_up[kMaxDevices][kUploadThreads];  // per device, created on first use, kept

// bytes [a, b) of the pieces taken end to end, copied on st
hipError_t upload_range(const HostPiece* pc, int np, size_t a, size_t b, hipStream_t st) {
  size_t base = 0;
  for (int i = 0; i < np; ++i) {
    const size_t lo = std::max(a, base), hi = std::min(b, base + pc[i].bytes);
    if (lo < hi) {
      const hipError_t e = hipMemcpyAsync(static_cast<char*>(pc[i].dst) + (lo - base),
                                          static_cast<const char*>(pc[i].src) + (lo - base), hi - lo,
                                          hipMemcpyHostToDevice, st);
      if (e != hipSuccess) return e;
    }
    base += pc[i].bytes;
  }
  return hipStreamSynchronize(st);
}

}  // namespace

void upload_pieces(const HostPiece* pc, int np, hipStream_t s) {
  size_t total = 0;
  for (int i = 0; i < np; ++i) total += pc[i].bytes;
  // the destinations may be blocks that earlier work on s still uses
  HIP_CHECK(hipStreamSynchronize(s));
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) fail(FEM_ERR_CAPACITY, "device index beyond the upload stream table");
  hipStream_t st[kUploadThreads];
  {
    // the streams are created, and each takes a first small copy (a prefix
    // of the first piece, so what lands there is the data itself), on the
    // first upload of the process, whatever its size: a stream's first copy
    // costs milliseconds more than the later ones
    std::lock_guard<std::mutex> lk(g_up_mu);
    for (int k = 0; k < kUploadThreads; ++k) {
      if (!g_up[dev][k]) {
        HIP_CHECK(hipStreamCreateWithFlags(&g_up[dev][k], hipStreamNonBlocking));
        const HostPiece* first = nullptr;
        for (int i = 0; i < np && !first; ++i)
          if (pc[i].bytes) first = &pc[i];
        if (first) {
          HIP_CHECK(hipMemcpyAsync(first->dst, first->src, std::min<size_t>(first->bytes, 4096),
                                   hipMemcpyHostToDevice, g_up[dev][k]));
          HIP_CHECK(hipStreamSynchronize(g_up[dev][k]));
        }
      }
      st[k] = g_up[dev][k];
    }
  }
  if (total < kUploadSplit) {
    for (int i = 0; i < np; ++i)
      if (pc[i].bytes) HIP_CHECK(hipMemcpyAsync(pc[i].dst, pc[i].src, pc[i].bytes, hipMemcpyHostToDevice, s));
    return;
  }
  hipError_t err[kUploadThreads];
  std::vector<std::thread> th;
  for (int k = 0; k < kUploadThreads; ++k) {
    const size_t a = total * k / kUploadThreads, b = total * (k + 1) / kUploadThreads;
    auto work = [&err, pc, np, a, b, dev, k, &st]() {
      hipError_t e = hipSetDevice(dev);  // the current device is per thread
      err[k] = e == hipSuccess ? upload_range(pc, np, a, b, st[k]) : e;
    };
    try {
      th.emplace_back(work);
    } catch (...) {
      work();  // no thread: copy this share here
    }
  }
  for (auto& t : th) t.join();
  for (int k = 0; k < kUploadThreads; ++k) HIP_CHECK(err[k]);
}

}  // namespace fem
