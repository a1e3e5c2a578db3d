// Setup-phase device allocation pool (see DevPoolScope in common.hpp).

#include <map>
#include <unordered_map>

#include "common.hpp"

namespace fem {

namespace {

struct Pool {
  std::unordered_map<void*, size_t> owned;  // blocks this pool allocated -> size class
  std::multimap<size_t, void*> free_;       // released blocks by size class
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

}  // namespace

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
  HIP_CHECK(hipMalloc(&p, c));
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

DevPoolScope::DevPoolScope() : prev_(g_pool) { g_pool = new Pool(); }

DevPoolScope::~DevPoolScope() {
  Pool* pl = g_pool;
  for (auto& kv : pl->free_) (void)hipFree(kv.second);
  delete pl;
  g_pool = static_cast<Pool*>(prev_);
}

}  // namespace fem
