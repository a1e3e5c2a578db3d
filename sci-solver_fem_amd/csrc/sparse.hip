// Row-wise sparse products for the AMG setup (Gustavson).
//
// Replaces the CUSP expand-sort-reduce products of the reference
// (cusp::multiply at src/core/cuda/smoothedMG_amg_level.cu:394,397 and
// cusp::transpose at :382).
//
// Symbolic: one thread per output row inserts the row's column indices into
// a private open-addressing hash table in LDS (256 slots).  A row that does
// not fit is finished by a second kernel with a private table in global
// memory sized from its product count.  Columns are then sorted per row.
// Numeric: one thread per output row walks its products in a fixed order and
// accumulates into the row (LDS when it fits, otherwise its global slice),
// finding each column by binary search.  No float atomics: results are
// bit-reproducible.
#include <climits>

#include "fem_internal.hpp"
#include "kernels.hpp"
#include "half.hpp"

#include <algorithm>
#include <type_traits>

namespace fem {

namespace {

constexpr int kSymBlock = 64;
constexpr int kSymCap = 256;             // LDS hash slots per thread (power of two), long product rows
constexpr int kSymCapSmall = 128;        // the same for products of ~100 or fewer terms per row
constexpr double kStashBytes = 8e9;      // largest per-row column stash of the one-pass symbolic product

// A thread's table in LDS, lane-interleaved: entry t of thread j of a block
// of B threads is element t * B + j, so the lanes of a wavefront address
// distinct banks whichever entries they touch.  (With each thread's entries
// contiguous, a run of CAP = 32 to 256 words, entry t of every lane falls in
// one bank, and the scans and binary searches that visit the same t on all
// lanes serialise.)
template <typename T, int B>
struct LaneTab {
  T* p;
  __device__ __forceinline__ T& operator[](int t) const { return p[t * B]; }
};

__device__ __forceinline__ unsigned hash_slot(int key, unsigned mask) {
  return ((unsigned)key * 2654435761u) & mask;
}

// returns false when the table already holds `limit` keys and key is new
template <typename Tab>
__device__ __forceinline__ bool hash_insert(Tab tab, unsigned mask, int key, int& count, int limit) {
  unsigned p = hash_slot(key, mask);
  while (true) {
    int t = tab[p];
    if (t == key) return true;
    if (t == -1) {
      if (count >= limit) return false;
      tab[p] = key;
      ++count;
      return true;
    }
    p = (p + 1) & mask;
  }
}

// Enumerate the output columns of row i and insert them; false on overflow.
template <typename Tab>
__device__ __forceinline__ bool row_columns(int64_t i, const int32_t* __restrict__ l_rp, const int32_t* __restrict__ l_ci,
                                            const int32_t* __restrict__ r_rp, const int32_t* __restrict__ r_ci,
                                            const int32_t* __restrict__ colmap, bool drop_diag, Tab tab,
                                            unsigned mask, int& count, int limit) {
  for (int p = l_rp[i]; p < l_rp[i + 1]; ++p) {
    int k = l_ci[p];
    if (r_rp) {
      for (int q = r_rp[k]; q < r_rp[k + 1]; ++q) {
        int c = colmap ? colmap[r_ci[q]] : r_ci[q];
        if (drop_diag && c == (int)i) continue;
        if (!hash_insert(tab, mask, c, count, limit)) return false;
      }
    } else {
      int c = colmap ? colmap[k] : k;
      if (drop_diag && c == (int)i) continue;
      if (!hash_insert(tab, mask, c, count, limit)) return false;
    }
  }
  return true;
}

// move keys to the front of tab and sort them ascending (insertion sort)
template <typename Tab>
__device__ __forceinline__ void compact_sort_small(Tab tab, int cap, int count) {
  int m = 0;
  for (int t = 0; t < cap; ++t) {
    int v = tab[t];
    if (v != -1) tab[m++] = v;
  }
  for (int a = 1; a < count; ++a) {
    int v = tab[a];
    int b = a - 1;
    while (b >= 0 && tab[b] > v) { tab[b + 1] = tab[b]; --b; }
    tab[b + 1] = v;
  }
}

__device__ __forceinline__ void sift_down(int* a, int start, int end) {
  int root = start;
  while (2 * root + 1 < end) {
    int child = 2 * root + 1;
    if (child + 1 < end && a[child] < a[child + 1]) ++child;
    if (a[root] < a[child]) {
      int t = a[root]; a[root] = a[child]; a[child] = t;
      root = child;
    } else {
      return;
    }
  }
}

__device__ __forceinline__ void compact_sort_large(int* tab, int64_t cap, int count) {
  int m = 0;
  for (int64_t t = 0; t < cap; ++t) {
    int v = tab[t];
    if (v != -1) tab[m++] = v;
  }
  for (int st = count / 2 - 1; st >= 0; --st) sift_down(tab, st, count);
  for (int end = count - 1; end > 0; --end) {
    int t = tab[0]; tab[0] = tab[end]; tab[end] = t;
    sift_down(tab, 0, end);
  }
}

template <bool kFill, int CAP>
__global__ void __launch_bounds__(kSymBlock) sym_lds_kernel(int64_t nrows, const int32_t* __restrict__ l_rp,
                                                          const int32_t* __restrict__ l_ci,
                                                          const int32_t* __restrict__ r_rp,
                                                          const int32_t* __restrict__ r_ci,
                                                          const int32_t* __restrict__ colmap, int drop_diag,
                                                          int32_t* __restrict__ cnt, const int32_t* __restrict__ c_rp,
                                                          int32_t* __restrict__ c_ci) {
  __shared__ int lds[kSymBlock * CAP];
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrows) return;
  if (kFill && cnt[i] < 0) return;
  const LaneTab<int, kSymBlock> tab{lds + threadIdx.x};
  for (int t = 0; t < CAP; ++t) tab[t] = -1;
  int count = 0;
  bool ok = row_columns(i, l_rp, l_ci, r_rp, r_ci, colmap, drop_diag != 0, tab, CAP - 1, count, CAP * 3 / 4);
  if (!kFill) {
    cnt[i] = ok ? count : -1;
    return;
  }
  compact_sort_small(tab, CAP, count);
  int o = c_rp[i];
  for (int t = 0; t < count; ++t) c_ci[o + t] = tab[t];
}

// One pass for rows that fit an LDS table of CAP slots: G lanes per row
// insert the row's product columns into the row's hash table (LDS atomics:
// the lanes of a row split the entries of each R row, or the L entries when
// there is no R), then compact the table, rank-sort it and keep the row's
// sorted columns in a per-row stash of CAP * 3 / 4 slots.  The set of
// columns, its count and whether it exceeds the table's fill limit do not
// depend on the insertion order, so the result is deterministic; rows over
// the limit are marked with count -1 for the global path.  A table per row
// rather than per thread keeps G times fewer tables in LDS (one thread per
// row with 256 slots held a wavefront's 64 tables in 64 KB).
constexpr int kSymNT = 256;
template <int G, int CAP>
__global__ void __launch_bounds__(kSymNT) sym_group_stash_kernel(int64_t nrows, const int32_t* __restrict__ l_rp,
                                                                const int32_t* __restrict__ l_ci,
                                                                const int32_t* __restrict__ r_rp,
                                                                const int32_t* __restrict__ r_ci,
                                                                const int32_t* __restrict__ colmap, int drop_diag,
                                                                int32_t* __restrict__ cnt, int32_t* __restrict__ stash,
                                                                uint8_t* __restrict__ fit) {
  constexpr int RPB = kSymNT / G;
  constexpr int S = CAP + 1;  // padded row stride: the groups of a wavefront start on different banks
  constexpr int kLimit = CAP * 3 / 4;
  __shared__ int s_tab[RPB * S];
  __shared__ int s_cnt[RPB];
  __shared__ int s_over[RPB];
  const int g = (int)threadIdx.x / G, lane = (int)threadIdx.x % G;
  const int64_t i = (int64_t)xcd_block(blockIdx.x, gridDim.x) * RPB + g;  // XCD-contiguous rows
  const bool live = i < nrows;
  int* tab = s_tab + g * S;
  for (int t = lane; t < CAP; t += G) tab[t] = -1;
  if (lane == 0) {
    s_cnt[g] = 0;
    s_over[g] = 0;
  }
  __syncthreads();
  if (live) {
    auto insert = [&](int c) {
      if (drop_diag && c == (int)i) return;
      unsigned p = hash_slot(c, CAP - 1);
      for (int probe = 0; probe < CAP; ++probe) {
        const int old = atomicCAS(&tab[p], -1, c);
        if (old == c) return;
        if (old == -1) {
          if (atomicAdd(&s_cnt[g], 1) >= kLimit) s_over[g] = 1;
          return;
        }
        p = (p + 1) & (CAP - 1);
      }
      s_over[g] = 1;  // full table (more keys than the limit)
    };
    if (r_rp) {
      for (int p = l_rp[i]; p < l_rp[i + 1] && !s_over[g]; ++p) {
        const int k = l_ci[p];
        for (int q = r_rp[k] + lane; q < r_rp[k + 1]; q += G) insert(colmap ? colmap[r_ci[q]] : r_ci[q]);
      }
    } else {
      for (int p = l_rp[i] + lane; p < l_rp[i + 1] && !s_over[g]; p += G) {
        const int k = l_ci[p];
        insert(colmap ? colmap[k] : k);
      }
    }
  }
  __syncthreads();
  const bool ok = live && !s_over[g];
  const int count = s_cnt[g];
  if (live && lane == 0) {
    fit[i] = ok ? 1 : 0;
    cnt[i] = ok ? count : -1;
  }
  if (!ok) return;  // the whole group leaves
  // compact in place: a chunk's keys are all read (one load instruction of
  // the group) before any is written, and writes land at or below the chunk
  const int gbase = (int)(threadIdx.x & 63) & ~(G - 1);
  int base = 0;
  for (int t0 = 0; t0 < CAP; t0 += G) {
    const int key = tab[t0 + lane];
    const unsigned gm = (unsigned)((__ballot(key != -1) >> gbase) & ((1ull << G) - 1));
    if (key != -1) tab[base + __popc(gm & ((1u << lane) - 1))] = key;
    base += __popc(gm);
  }
  // rank sort (distinct keys)
  int32_t* o = stash + i * kLimit;
  for (int t = lane; t < count; t += G) {
    const int key = tab[t];
    int r = 0;
    for (int u = 0; u < count; ++u) r += tab[u] < key;
    o[r] = key;
  }
}

// the stashed rows into their CSR positions, kCopyLanes lanes per row
constexpr int kCopyLanes = 16;
__global__ void stash_copy_kernel(int64_t nrows, int stride, const uint8_t* __restrict__ fit,
                                  const int32_t* __restrict__ stash, const int32_t* __restrict__ c_rp,
                                  int32_t* __restrict__ c_ci) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = gid / kCopyLanes;
  const int t = (int)(gid % kCopyLanes);
  if (i >= nrows || !fit[i]) return;
  const int o = c_rp[i], len = c_rp[i + 1] - o;
  const int32_t* src = stash + i * stride;
  for (int q = t; q < len; q += kCopyLanes) c_ci[o + q] = src[q];
}

__global__ void collect_overflow_kernel(int64_t nrows, const int32_t* __restrict__ cnt, const int32_t* __restrict__ l_rp,
                                        const int32_t* __restrict__ l_ci, const int32_t* __restrict__ r_rp,
                                        int64_t ncols, int32_t* __restrict__ rows, int64_t* __restrict__ cap,
                                        unsigned* __restrict__ nover) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrows || cnt[i] >= 0) return;
  int64_t ub = 0;
  for (int p = l_rp[i]; p < l_rp[i + 1]; ++p) ub += r_rp ? (r_rp[l_ci[p] + 1] - r_rp[l_ci[p]]) : 1;
  if (ub > ncols) ub = ncols;
  int64_t c = 1;
  while (c < 2 * ub + 2) c <<= 1;
  unsigned slot = atomicAdd(nover, 1u);
  rows[slot] = (int32_t)i;
  cap[slot] = c;
}

template <bool kFill>
__global__ void sym_global_kernel(int64_t nover, const int32_t* __restrict__ rows, const int64_t* __restrict__ off,
                                  const int32_t* __restrict__ l_rp, const int32_t* __restrict__ l_ci,
                                  const int32_t* __restrict__ r_rp, const int32_t* __restrict__ r_ci,
                                  const int32_t* __restrict__ colmap, int drop_diag, int* __restrict__ scratch,
                                  int32_t* __restrict__ cnt, const int32_t* __restrict__ c_rp,
                                  int32_t* __restrict__ c_ci) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nover) return;
  int64_t i = rows[t];
  int* tab = scratch + off[t];
  int64_t cap = off[t + 1] - off[t];
  for (int64_t q = 0; q < cap; ++q) tab[q] = -1;
  int count = 0;
  row_columns(i, l_rp, l_ci, r_rp, r_ci, colmap, drop_diag != 0, tab, (unsigned)(cap - 1), count, 0x7fffffff);
  if (!kFill) {
    cnt[i] = count;
    return;
  }
  compact_sort_large(tab, cap, count);
  int o = c_rp[i];
  for (int q = 0; q < count; ++q) c_ci[o + q] = tab[q];
}

template <typename C>
__device__ __forceinline__ int bsearch_col(C cols, int len, int c) {
  int lo = 0, hi = len;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (cols[mid] < c) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Numeric product, G lanes per row: the row's columns and accumulators in
// LDS when the row has at most CAP columns.  The L entries of the row are
// taken in order; for each, the lanes split the entries of the matching R
// row, which fall on distinct output columns, so the lanes add into distinct
// slots and every output value takes its terms in (L entry, R entry) order:
// the order, and the result bit for bit, of one thread walking the row.
// LDS ordering across the L entries: the group's lanes are in one wavefront,
// whose LDS accesses execute in program order.  Rows longer than CAP are
// summed in the output row itself by the group's first lane (global memory
// gives no such order between lanes).  A table per row rather than per
// thread keeps G times fewer tables in LDS: with one thread per row and 64
// slots, 48 KB per wavefront left 3 wavefronts per CU.
constexpr int kNumNT = 256;
template <int G, int CAP>
__global__ void __launch_bounds__(kNumNT) num_group_kernel(int64_t nrows, const int32_t* __restrict__ l_rp,
                                                          const int32_t* __restrict__ l_ci,
                                                          const double* __restrict__ l_v,
                                                          const int32_t* __restrict__ r_rp,
                                                          const int32_t* __restrict__ r_ci,
                                                          const double* __restrict__ r_v,
                                                          const int32_t* __restrict__ c_rp,
                                                          const int32_t* __restrict__ c_ci, double* __restrict__ c_v) {
  constexpr int RPB = kNumNT / G;
  constexpr int S = CAP + 1;  // padded row stride: the groups of a wavefront start on different banks
  __shared__ int s_cols[RPB * S];
  __shared__ double s_acc[RPB * S];
  const int g = (int)threadIdx.x / G, lane = (int)threadIdx.x % G;
  // XCD-contiguous rows: neighbouring rows gather the same R rows through one L2
  const int64_t i = (int64_t)xcd_block(blockIdx.x, gridDim.x) * RPB + g;
  const bool live = i < nrows;
  const int o = live ? c_rp[i] : 0;
  const int len = live ? c_rp[i + 1] - o : 0;
  const bool in_lds = len <= CAP;
  int* cols = s_cols + g * S;
  double* acc = s_acc + g * S;
  if (in_lds)
    for (int t = lane; t < len; t += G) {
      cols[t] = c_ci[o + t];
      acc[t] = 0.0;
    }
  __syncthreads();
  if (live && in_lds) {
    for (int p = l_rp[i]; p < l_rp[i + 1]; ++p) {
      const int k = l_ci[p];
      const double a = l_v[p];
      for (int q = r_rp[k] + lane; q < r_rp[k + 1]; q += G) acc[bsearch_col(cols, len, r_ci[q])] += a * r_v[q];
    }
  }
  __syncthreads();
  if (in_lds)
    for (int t = lane; t < len; t += G) c_v[o + t] = acc[t];
  if (live && !in_lds && lane == 0) {
    const int32_t* gc = c_ci + o;
    double* ga = c_v + o;
    for (int t = 0; t < len; ++t) ga[t] = 0.0;
    for (int p = l_rp[i]; p < l_rp[i + 1]; ++p) {
      const int k = l_ci[p];
      const double a = l_v[p];
      for (int q = r_rp[k]; q < r_rp[k + 1]; ++q) ga[bsearch_col(gc, len, r_ci[q])] += a * r_v[q];
    }
  }
}

// longest row of a CSR matrix (a capped grid, one atomic per block)
__global__ void __launch_bounds__(256) max_row_len_kernel(int64_t n, const int32_t* __restrict__ rp,
                                                        int* __restrict__ out) {
  __shared__ int s_part[4];
  int m = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    m = max(m, rp[i + 1] - rp[i]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int t = max(max(s_part[0], s_part[1]), max(s_part[2], s_part[3]));
    if (t) atomicMax(out, t);
  }
}

// Coarse-level products (few rows, thousands of products per row): a
// wavefront per row.  Symbolic: the row's columns are marked in an LDS bitmap
// of ncols bits (a set, so the marking order does not matter) and read back
// in ascending order.  Numeric: the L entries of the row are taken in order;
// for each, the lanes split the entries of the matching R row, which hit
// distinct output columns, so every output value is summed in the same order
// as in num_group_kernel and the result is bitwise the same.
constexpr int kWaveBitmapCols = 1 << 17;  // 16 KB bitmap per row
constexpr int kWaveAccCap = 2048;         // LDS accumulators per row (16 KB)

template <bool kFill>
__global__ void __launch_bounds__(64) sym_wave_kernel(int64_t nrows, const int32_t* __restrict__ l_rp,
                                                    const int32_t* __restrict__ l_ci,
                                                    const int32_t* __restrict__ r_rp,
                                                    const int32_t* __restrict__ r_ci,
                                                    const int32_t* __restrict__ colmap, int drop_diag, int64_t ncols,
                                                    int32_t* __restrict__ cnt, const int32_t* __restrict__ c_rp,
                                                    int32_t* __restrict__ c_ci) {
  extern __shared__ unsigned bits[];
  const int64_t i = blockIdx.x;
  const int lane = threadIdx.x;
  const int nw = (int)((ncols + 31) >> 5);
  for (int w = lane; w < nw; w += 64) bits[w] = 0u;
  __syncthreads();
  for (int p = l_rp[i]; p < l_rp[i + 1]; ++p) {
    const int k = l_ci[p];
    if (r_rp) {
      for (int q = r_rp[k] + lane; q < r_rp[k + 1]; q += 64) {
        const int c = colmap ? colmap[r_ci[q]] : r_ci[q];
        if (!(drop_diag && c == (int)i)) atomicOr(&bits[c >> 5], 1u << (c & 31));
      }
    } else if (lane == 0) {
      const int c = colmap ? colmap[k] : k;
      if (!(drop_diag && c == (int)i)) atomicOr(&bits[c >> 5], 1u << (c & 31));
    }
  }
  __syncthreads();
  // lane owns words [w0, w1); exclusive prefix of their popcounts gives its output offset
  const int per = (nw + 63) / 64;
  const int w0 = min(nw, lane * per), w1 = min(nw, w0 + per);
  int mine = 0;
  for (int w = w0; w < w1; ++w) mine += __popc(bits[w]);
  int incl = mine;
  for (int off = 1; off < 64; off <<= 1) {
    const int v = __shfl_up(incl, off, 64);
    if (lane >= off) incl += v;
  }
  const int total = __shfl(incl, 63, 64);
  if (!kFill) {
    if (lane == 0) cnt[i] = total;
    return;
  }
  int o = c_rp[i] + incl - mine;
  for (int w = w0; w < w1; ++w) {
    unsigned b = bits[w];
    while (b) {
      const int t = __ffs(b) - 1;
      c_ci[o++] = (w << 5) + t;
      b &= b - 1u;
    }
  }
}

__global__ void __launch_bounds__(64) num_wave_kernel(int64_t nrows, const int32_t* __restrict__ l_rp,
                                                    const int32_t* __restrict__ l_ci,
                                                    const double* __restrict__ l_v,
                                                    const int32_t* __restrict__ r_rp,
                                                    const int32_t* __restrict__ r_ci,
                                                    const double* __restrict__ r_v,
                                                    const int32_t* __restrict__ c_rp,
                                                    const int32_t* __restrict__ c_ci, double* __restrict__ c_v) {
  __shared__ double s_acc[kWaveAccCap];
  const int64_t i = blockIdx.x;
  const int lane = threadIdx.x;
  const int o = c_rp[i];
  const int len = c_rp[i + 1] - o;
  const int32_t* cols = c_ci + o;
  double* acc = len <= kWaveAccCap ? s_acc : c_v + o;
  for (int t = lane; t < len; t += 64) acc[t] = 0.0;
  __syncthreads();
  for (int p = l_rp[i]; p < l_rp[i + 1]; ++p) {
    const int k = l_ci[p];
    const double a = l_v[p];
    for (int q = r_rp[k] + lane; q < r_rp[k + 1]; q += 64) acc[bsearch_col(cols, len, r_ci[q])] += a * r_v[q];
    __syncthreads();  // the next L entry may hit the same output columns
  }
  if (len <= kWaveAccCap)
    for (int t = lane; t < len; t += 64) c_v[o + t] = acc[t];
}

__global__ void transpose_keys_kernel(int64_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                      uint64_t* __restrict__ key, int32_t* __restrict__ idx) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int p = rp[i]; p < rp[i + 1]; ++p) {
    key[p] = ((uint64_t)(uint32_t)ci[p] << 32) | (uint64_t)(uint32_t)i;
    idx[p] = p;
  }
}

__global__ void transpose_fill_kernel(int64_t nnz, const uint64_t* __restrict__ key, const int32_t* __restrict__ idx,
                                      const double* __restrict__ v, int32_t* __restrict__ ci, double* __restrict__ rv) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nnz) return;
  ci[t] = (int32_t)(key[t] & 0xffffffffull);
  rv[t] = v[idx[t]];
}

__global__ void transpose_ptr_kernel(int64_t m, const uint64_t* __restrict__ key, int64_t nnz, int32_t* __restrict__ rp) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c > m) return;
  int64_t lo = 0, hi = nnz;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if ((int64_t)(key[mid] >> 32) < c) lo = mid + 1; else hi = mid;
  }
  rp[c] = (int32_t)lo;
}

// row stored at sliced-ELL position p (rows may be reordered by length)
__device__ __forceinline__ int64_t sell_row(const int32_t* __restrict__ srow, int64_t p) {
  return srow ? (int64_t)srow[p] : p;
}

// slices of 64 / g rows, g lanes per row: w[sl] = 64 * (steps of the longest row)
__global__ void sell_width_kernel(int64_t n, int64_t nslices, int g, const int32_t* __restrict__ rp,
                                  const int32_t* __restrict__ srow, int32_t* __restrict__ w) {
  int64_t sl = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (sl >= nslices) return;
  const int rows = 64 / g;
  int mx = 0;
  for (int64_t p = sl * rows; p < sl * rows + rows && p < n; ++p) {
    const int64_t r = sell_row(srow, p);
    mx = max(mx, (rp[r + 1] - rp[r] + g - 1) / g);
  }
  w[sl] = mx * 64;
}

// sort key of the length-ordered layout: window of `win` rows, then longest first
__global__ void sell_sort_keys_kernel(int64_t n, const int32_t* __restrict__ rp, int64_t win,
                                      uint32_t* __restrict__ key, int32_t* __restrict__ idx) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int len = min(rp[r + 1] - rp[r], 255);
  key[r] = ((uint32_t)(r / win) << 8) | (uint32_t)(255 - len);
  idx[r] = (int32_t)r;
}

// Window bases of each slice, first pass, a wavefront per slice (lane = row
// position): when the slice's columns span fewer than kSellWindows aligned
// blocks (col >> kSellWinBits) from its lowest one, those consecutive blocks
// are its windows (compact = 1); otherwise compact = -1 and the second pass
// decides.  A slice without entries is not compact (0).
__global__ void sell_window_range_kernel(int64_t n, int64_t nslices, int g, const int32_t* __restrict__ rp,
                                         const int32_t* __restrict__ ci, const int32_t* __restrict__ srow,
                                         int32_t* __restrict__ swin, int32_t* __restrict__ compact) {
  const int64_t sl = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (sl >= nslices) return;  // the whole wavefront
  const int rows = 64 / g;
  int bmin = 0x7fffffff, bmax = -1;
  const int64_t p = sl * rows + lane;
  if (lane < rows && p < n) {
    const int64_t r = sell_row(srow, p);
    for (int q = rp[r]; q < rp[r + 1]; ++q) {
      const int b = ci[q] >> kSellWinBits;
      bmin = min(bmin, b);
      bmax = max(bmax, b);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    bmin = min(bmin, __shfl_xor(bmin, off, 64));
    bmax = max(bmax, __shfl_xor(bmax, off, 64));
  }
  const bool fast = bmax >= 0 && bmax - bmin < kSellWindows;
  if (lane < kSellWindows) swin[sl * kSellWindows + lane] = fast ? (bmin + lane) << kSellWinBits : -1;
  if (lane == 0) compact[sl] = fast ? 1 : (bmax >= 0 ? -1 : 0);
}

// Second pass, a wavefront per slice left undecided (lane = row position):
// each lane collects its row's distinct blocks, then the wavefront takes the
// union in ascending order, one wave-wide minimum per block; more than
// kSellWindows of them: not compact.
__global__ void sell_window_kernel(int64_t n, int64_t nslices, int g, const int32_t* __restrict__ rp,
                                   const int32_t* __restrict__ ci, const int32_t* __restrict__ srow,
                                   int32_t* __restrict__ swin, int32_t* __restrict__ compact) {
  const int64_t sl = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (sl >= nslices || compact[sl] != -1) return;  // the whole wavefront
  const int rows = 64 / g;
  int mine[kSellWindows + 1];
  int nm = 0;
  bool over = false;  // this row alone has more than kSellWindows blocks
  const int64_t p = sl * rows + lane;
  if (lane < rows && p < n) {
    const int64_t r = sell_row(srow, p);
    for (int q = rp[r]; q < rp[r + 1] && !over; ++q) {
      const int b = ci[q] >> kSellWinBits;
      bool seen = false;
      for (int t = 0; t < nm; ++t) seen = seen || mine[t] == b;
      if (seen) continue;
      if (nm == kSellWindows) over = true;
      else mine[nm++] = b;
    }
  }
  bool ok = !__any(over);
  unsigned taken = 0;  // bit t: mine[t] is in the union
  int win[kSellWindows];
  int nb = 0;
  while (ok) {
    int cand = 0x7fffffff;
    for (int t = 0; t < nm; ++t)
      if (!(taken >> t & 1u)) cand = min(cand, mine[t]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) cand = min(cand, __shfl_xor(cand, off, 64));
    if (cand == 0x7fffffff) break;
    if (nb == kSellWindows) {
      ok = false;
      break;
    }
    win[nb++] = cand;
    for (int t = 0; t < nm; ++t)
      if (mine[t] == cand) taken |= 1u << t;
  }
  ok = ok && nb > 0;
  if (lane < kSellWindows) {
    int w = -1;
    if (ok) {
      w = win[0];
      for (int t = 0; t < kSellWindows; ++t)
        if (t == lane && t < nb) w = win[t];
      w <<= kSellWinBits;
    }
    swin[sl * kSellWindows + lane] = w;
  }
  if (lane == 0) compact[sl] = ok ? 1 : 0;
}

// stored nonzeros of each compact slice (0 for the others)
__global__ void sell_compact_nnz_kernel(int64_t n, int64_t nslices, int g, const int32_t* __restrict__ rp,
                                        const int32_t* __restrict__ srow, const int32_t* __restrict__ compact,
                                        int64_t* __restrict__ out) {
  int64_t sl = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (sl >= nslices) return;
  const int rows = 64 / g;
  int64_t cnt = 0;
  if (compact[sl])
    for (int64_t p = sl * rows; p < sl * rows + rows && p < n; ++p) {
      const int64_t r = sell_row(srow, p);
      cnt += rp[r + 1] - rp[r];
    }
  out[sl] = cnt;
}

// CSR entries of a slice staged in LDS by sell_fill_lds_kernel (slices of
// one lane per row in row order whose rows hold at most this many entries)
constexpr int kFillSpan = 1024;

// the slice's span of CSR entries when its 64 rows are consecutive
__device__ __forceinline__ int fill_span(int64_t n, int64_t sl, const int32_t* __restrict__ rp) {
  const int64_t r0 = sl * 64, r1 = r0 + 64 < n ? r0 + 64 : n;
  return r0 < n ? rp[r1] - rp[r0] : 0;
}

// one thread per (slice, lane): position = slice * (64 / g) + lane / g, and
// the lane holds entries lane % g, lane % g + g, ... of that row.  With
// skip_staged, slices that sell_fill_lds_kernel fills are left to it.
__global__ void sell_fill_kernel(int64_t n, int64_t nslices, int g, const int32_t* __restrict__ rp,
                                 const int32_t* __restrict__ ci, const double* __restrict__ v,
                                 const int32_t* __restrict__ srow, const int32_t* __restrict__ sp,
                                 const int32_t* __restrict__ swin, int32_t* __restrict__ sc,
                                 uint16_t* __restrict__ sc16, double* __restrict__ s64, float* __restrict__ s32,
                                 uint16_t* __restrict__ sh, int skip_staged) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nslices * 64) return;
  int64_t sl = t >> 6;
  int lane = (int)(t & 63);
  if (skip_staged && fill_span(n, sl, rp) <= kFillSpan) return;
  const int sub = lane % g;
  const int64_t pos = sl * (64 / g) + lane / g;
  int64_t row = pos < n ? sell_row(srow, pos) : pos;
  int off = sp[sl];
  int w = (sp[sl + 1] - off) >> 6;
  int len = 0, base = 0;
  if (row < n) {
    base = rp[row];
    len = rp[row + 1] - base;
  }
  const int32_t* win = swin + sl * kSellWindows;
  const bool compact = win[0] >= 0;
  // padding repeats the row's last column (a valid index for rectangular P
  // too, and already in cache); empty rows and rows past n use the slice's
  // first window base (a column at or below one the slice holds) or column 0
  int pad_col = len > 0 ? ci[base + len - 1] : (compact ? win[0] : 0);
  for (int k = 0; k < w; ++k) {
    int64_t e = off + (int64_t)k * 64 + lane;
    const int j = k * g + sub;
    const int col = j < len ? ci[base + j] : pad_col;
    sc[e] = col;
    if (compact) {
      int slot = 0;
      for (int i = 0; i < kSellWindows; ++i)
        if (win[i] == ((col >> kSellWinBits) << kSellWinBits)) {
          slot = i;
          break;
        }
      sc16[e] = (uint16_t)((slot << kSellWinBits) | (col - win[slot]));
    } else {
      sc16[e] = 0;
    }
    const double val = j < len ? v[base + j] : 0.0;
    if (s64) s64[e] = val;
    if (s32) s32[e] = (float)val;
    if (sh) sh[e] = half_bits(val);
  }
}

// sell_fill_kernel for one lane per row in row order, one wavefront per
// slice: the slice's rows are consecutive in the CSR arrays, so the
// wavefront reads their columns and values as one coalesced run into LDS;
// each lane then writes its row's entries in sliced order (the same values
// the generic kernel writes).  Slices spanning more than kFillSpan entries
// are left to sell_fill_kernel.
__global__ void __launch_bounds__(64) sell_fill_lds_kernel(int64_t n, int64_t nslices, const int32_t* __restrict__ rp,
                                                           const int32_t* __restrict__ ci,
                                                           const double* __restrict__ v,
                                                           const int32_t* __restrict__ sp,
                                                           const int32_t* __restrict__ swin,
                                                           int32_t* __restrict__ sc, uint16_t* __restrict__ sc16,
                                                           double* __restrict__ s64, float* __restrict__ s32,
                                                           uint16_t* __restrict__ sh) {
  __shared__ int32_t lci[kFillSpan];
  __shared__ double lv[kFillSpan];
  const int64_t sl = blockIdx.x;
  if (sl >= nslices) return;
  const int span = fill_span(n, sl, rp);
  if (span > kFillSpan) return;
  const int lane = threadIdx.x;
  const int64_t r0 = sl * 64;
  const int b0 = r0 < n ? rp[r0] : 0;
  for (int t = lane; t < span; t += 64) {
    lci[t] = ci[b0 + t];
    lv[t] = v[b0 + t];
  }
  __syncthreads();
  const int64_t row = r0 + lane;
  int len = 0, base = 0;
  if (row < n) {
    base = rp[row] - b0;
    len = rp[row + 1] - rp[row];
  }
  const int off = sp[sl];
  const int w = (sp[sl + 1] - off) >> 6;
  const int32_t* win = swin + sl * kSellWindows;
  const bool compact = win[0] >= 0;
  const int pad_col = len > 0 ? lci[base + len - 1] : (compact ? win[0] : 0);
  for (int k = 0; k < w; ++k) {
    const int64_t e = off + (int64_t)k * 64 + lane;
    const int col = k < len ? lci[base + k] : pad_col;
    sc[e] = col;
    if (compact) {
      int slot = 0;
      for (int i = 0; i < kSellWindows; ++i)
        if (win[i] == ((col >> kSellWinBits) << kSellWinBits)) {
          slot = i;
          break;
        }
      sc16[e] = (uint16_t)((slot << kSellWinBits) | (col - win[slot]));
    } else {
      sc16[e] = 0;
    }
    const double val = k < len ? lv[base + k] : 0.0;
    if (s64) s64[e] = val;
    if (s32) s32[e] = (float)val;
    if (sh) sh[e] = half_bits(val);
  }
}

// Off-diagonal entries per row of a square CSR matrix (count pass of the split
// into the off-diagonal part and the diagonal).
__global__ void split_diag_count_kernel(int64_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                        int32_t* __restrict__ cnt) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int c = 0;
  for (int q = rp[i]; q < rp[i + 1]; ++q) c += ci[q] != (int)i;
  cnt[i] = c;
}

// Fill pass of the split with kSplitLanes lanes per row walking the row in
// chunks: an entry's position among the off-diagonal ones is its position in
// the row less the diagonal entries before it (a ballot per chunk), so the
// lanes write consecutive addresses; the diagonal value is the last diagonal
// entry's, as in the count-and-fill kernel above (0 where a row has none).
// One thread per row writing its own run took 2.6 ms at 4.2M rows.
constexpr int kSplitLanes = 16;
__global__ void __launch_bounds__(256) split_diag_fill_kernel(int64_t n, const int32_t* __restrict__ rp,
                                                            const int32_t* __restrict__ ci,
                                                            const double* __restrict__ v,
                                                            const int32_t* __restrict__ orp,
                                                            int32_t* __restrict__ oci, double* __restrict__ ov,
                                                            double* __restrict__ diag) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = gid / kSplitLanes;
  const int t = (int)(gid % kSplitLanes);
  if (i >= n) return;  // the lanes of a row leave together
  const int gbase = (int)(threadIdx.x & 63) & ~(kSplitLanes - 1);
  const int r0 = rp[i], r1 = rp[i + 1];
  const int o = orp[i];
  int before = 0;  // diagonal entries in the earlier chunks
  int dq = -1;     // this lane's last diagonal entry
  double dv = 0.0;
  for (int q0 = r0; q0 < r1; q0 += kSplitLanes) {
    const int q = q0 + t;
    const bool in = q < r1;
    const int j = in ? ci[q] : -1;
    const bool isd = in && j == (int)i;
    const unsigned gm = (unsigned)((__ballot(isd) >> gbase) & ((1ull << kSplitLanes) - 1));
    if (isd) {
      dq = q;
      dv = v[q];
    } else if (in) {
      const int dst = o + (q - r0) - before - __popc(gm & ((1u << t) - 1));
      oci[dst] = j;
      ov[dst] = v[q];
    }
    before += __popc(gm);
  }
#pragma unroll
  for (int off = kSplitLanes / 2; off > 0; off >>= 1) {
    const int oq = __shfl_xor(dq, off, kSplitLanes);
    const double odv = __shfl_xor(dv, off, kSplitLanes);
    if (oq > dq) {
      dq = oq;
      dv = odv;
    }
  }
  if (t == 0) diag[i] = dv;
}

// ---- LDS staging lists -----------------------------------------------------------

constexpr int kLdsBitWords = 8192;  // bitmap words of one block: a column range of 262144
constexpr int kLdsHashBits = 13;
constexpr int kLdsHash = 1 << kLdsHashBits;  // hash slots of a wider block
constexpr int kLdsMaxCols = 8192;   // distinct columns of one block staged at solve time
constexpr int kLdsThreads = 256;

// Block b (slices [b bs, (b+1) bs)): lo[b] = its lowest stored column, nw[b] =
// the 32-bit words that a bitmap of its column range needs.  flag[0] = max nw
// up to kLdsBitWords; blocks above it (a Morton order's jumps) are appended to
// wide[0..flag[2]) and set nw[b] = -1.
__global__ void __launch_bounds__(kLdsThreads) lds_range_kernel(int64_t nslices, int bs, const int32_t* __restrict__ sp,
                                                              const int32_t* __restrict__ sc, int32_t* __restrict__ lo,
                                                              int32_t* __restrict__ nw, int* __restrict__ flag,
                                                              int32_t* __restrict__ wide) {
  __shared__ int red[2][kLdsThreads / 64];
  const int64_t b = blockIdx.x;
  const int64_t s0 = b * bs, s1 = min(nslices, s0 + (int64_t)bs);
  const int e0 = sp[s0], e1 = sp[s1];
  int mn = INT_MAX, mx = -1;
  for (int e = e0 + (int)threadIdx.x; e < e1; e += kLdsThreads) {
    const int c = sc[e];
    mn = min(mn, c);
    mx = max(mx, c);
  }
  for (int off = 32; off > 0; off >>= 1) {
    mn = min(mn, __shfl_xor(mn, off, 64));
    mx = max(mx, __shfl_xor(mx, off, 64));
  }
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][wave] = mn;
    red[1][wave] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kLdsThreads / 64; ++w) {
      mn = min(mn, red[0][w]);
      mx = max(mx, red[1][w]);
    }
    const int words = mx >= mn ? ((mx - mn) >> 5) + 1 : 0;
    lo[b] = mx >= mn ? mn : 0;
    if (words <= kLdsBitWords) {
      nw[b] = words;
      atomicMax(flag, words);
    } else {
      nw[b] = -1;
      wide[atomicAdd(flag + 2, 1)] = (int32_t)b;
    }
  }
}

// Block b marks its stored columns in an LDS bitmap over [lo[b], lo[b] + 32 nw[b]);
// the exclusive prefix of the words' popcounts ranks each column among the
// block's distinct columns, which come out ascending.  Count pass: cnt[b] =
// distinct columns; fill pass: the list at lcs[b] and each entry's rank.
template <bool kFill>
__global__ void __launch_bounds__(kLdsThreads) lds_cols_kernel(int64_t nslices, int bs, const int32_t* __restrict__ sp,
                                                             const int32_t* __restrict__ sc,
                                                             const int32_t* __restrict__ lo,
                                                             const int32_t* __restrict__ nwb, int32_t* __restrict__ cnt,
                                                             const int32_t* __restrict__ lcs,
                                                             int32_t* __restrict__ lcols, uint16_t* __restrict__ lc) {
  extern __shared__ uint32_t bits[];  // nw words of bitmap, then nw word prefixes
  __shared__ int wsum[kLdsThreads / 64];
  const int64_t b = blockIdx.x;
  const int64_t s0 = b * bs, s1 = min(nslices, s0 + (int64_t)bs);
  const int e0 = sp[s0], e1 = sp[s1];
  const int c0 = lo[b], nw = nwb[b];
  if (nw < 0) return;  // a wide block: lds_hash_kernel
  uint32_t* pre = bits + nw;
  for (int w = threadIdx.x; w < nw; w += kLdsThreads) bits[w] = 0u;
  __syncthreads();
  for (int e = e0 + (int)threadIdx.x; e < e1; e += kLdsThreads) {
    const int v = sc[e] - c0;
    atomicOr(&bits[v >> 5], 1u << (v & 31));
  }
  __syncthreads();
  // thread t owns words [w0, w1)
  const int per = (nw + kLdsThreads - 1) / kLdsThreads;
  const int w0 = min(nw, (int)threadIdx.x * per), w1 = min(nw, w0 + per);
  int mine = 0;
  for (int w = w0; w < w1; ++w) mine += __popc(bits[w]);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int incl = mine;
  for (int off = 1; off < 64; off <<= 1) {
    const int v = __shfl_up(incl, off, 64);
    if (lane >= off) incl += v;
  }
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  int base = 0, total = 0;
  for (int w = 0; w < kLdsThreads / 64; ++w) {
    if (w < wave) base += wsum[w];
    total += wsum[w];
  }
  if (!kFill) {
    if (threadIdx.x == 0) cnt[b] = total;
    return;
  }
  int r = base + incl - mine;
  const int o = lcs[b];
  for (int w = w0; w < w1; ++w) {
    pre[w] = (uint32_t)r;
    uint32_t m = bits[w];
    while (m) {
      lcols[o + r++] = c0 + (w << 5) + (__ffs(m) - 1);
      m &= m - 1u;
    }
  }
  __syncthreads();
  for (int e = e0 + (int)threadIdx.x; e < e1; e += kLdsThreads) {
    const int v = sc[e] - c0;
    const int w = v >> 5;
    lc[e] = (uint16_t)(pre[w] + __popc(bits[w] & ((1u << (v & 31)) - 1u)));
  }
}

// ascending bitonic sort of a[0..P), P a power of two, by the whole block
__device__ void block_bitonic_sort(int32_t* a, int P) {
  for (int k = 2; k <= P; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < P; i += kLdsThreads) {
        const int l = i ^ j;
        if (l > i) {
          const int x = a[i], y = a[l];
          const bool up = (i & k) == 0;
          if ((x > y) == up) {
            a[i] = y;
            a[l] = x;
          }
        }
      }
      __syncthreads();
    }
  }
}

// Wide block b = wide[blockIdx.x]: its stored columns go into an LDS hash set
// (linear probing, at most kLdsHash probes), the distinct ones are compacted
// and sorted, and each entry's index is found by binary search.  Count pass:
// cnt[b] = distinct columns; fill pass: the list at lcs[b] and each entry's
// index.  A set that overflows sets *over (and cnt[b] = 0).
template <bool kFill>
__global__ void __launch_bounds__(kLdsThreads) lds_hash_kernel(int64_t nslices, int bs, const int32_t* __restrict__ wide,
                                                             const int32_t* __restrict__ sp,
                                                             const int32_t* __restrict__ sc, int32_t* __restrict__ cnt,
                                                             const int32_t* __restrict__ lcs,
                                                             int32_t* __restrict__ lcols, uint16_t* __restrict__ lc,
                                                             int* __restrict__ over) {
  extern __shared__ int32_t sm[];  // kLdsHash slots, then kLdsMaxCols keys
  int32_t* table = sm;
  int32_t* keys = sm + kLdsHash;
  __shared__ int nu_s, bad;
  const int64_t b = wide[blockIdx.x];
  const int64_t s0 = b * bs, s1 = min(nslices, s0 + (int64_t)bs);
  const int e0 = sp[s0], e1 = sp[s1];
  for (int t = threadIdx.x; t < kLdsHash; t += kLdsThreads) table[t] = -1;
  if (threadIdx.x == 0) {
    nu_s = 0;
    bad = 0;
  }
  __syncthreads();
  for (int e = e0 + (int)threadIdx.x; e < e1; e += kLdsThreads) {
    const int c = sc[e];
    unsigned h = ((unsigned)c * 2654435761u) >> (32 - kLdsHashBits);
    int probe = 0;
    for (; probe < kLdsHash; ++probe) {
      const int prev = atomicCAS(&table[h], -1, c);
      if (prev == -1 || prev == c) break;
      h = (h + 1u) & (unsigned)(kLdsHash - 1);
    }
    if (probe == kLdsHash) bad = 1;
  }
  __syncthreads();
  if (bad) {
    if (threadIdx.x == 0) {
      if (!kFill) cnt[b] = 0;
      over[0] = 1;
    }
    return;
  }
  for (int t = threadIdx.x; t < kLdsHash; t += kLdsThreads) {
    const int c = table[t];
    if (c >= 0) keys[atomicAdd(&nu_s, 1)] = c;
  }
  __syncthreads();
  const int nu = nu_s;
  if (!kFill) {
    if (threadIdx.x == 0) cnt[b] = nu;
    return;
  }
  int P = 1;
  while (P < nu) P <<= 1;
  for (int t = nu + (int)threadIdx.x; t < P; t += kLdsThreads) keys[t] = INT_MAX;
  __syncthreads();
  block_bitonic_sort(keys, P);
  const int o = lcs[b];
  for (int t = threadIdx.x; t < nu; t += kLdsThreads) lcols[o + t] = keys[t];
  for (int e = e0 + (int)threadIdx.x; e < e1; e += kLdsThreads) {
    const int v = sc[e];
    int lo = 0, hi = nu;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (keys[mid] < v) lo = mid + 1; else hi = mid;
    }
    lc[e] = (uint16_t)lo;
  }
}

__global__ void max_i32_kernel(int64_t n, const int32_t* __restrict__ a, int* __restrict__ mx) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) atomicMax(mx, a[i]);
}

int bits_for(int64_t v) {
  int b = 1;
  while (b < 63 && ((int64_t)1 << b) <= v) ++b;
  return b;
}

}  // namespace

// coarse products: few rows, output columns that fit an LDS bitmap
bool use_wave_rows(int64_t nrows, int64_t ncols) { return nrows > 0 && nrows <= 65536 && ncols <= kWaveBitmapCols; }

void spgemm_symbolic(const Csr& L, const int32_t* r_rp, const int32_t* r_ci, int64_t r_rows, const int32_t* colmap,
                     int64_t ncols, bool drop_diag, Csr& C, hipStream_t s) {
  const int64_t n = L.n;
  C.n = n;
  C.m = ncols;
  C.rp.resize(n + 1);
  if (use_wave_rows(n, ncols)) {
    const size_t lds = sizeof(unsigned) * (size_t)((ncols + 31) / 32 > 0 ? (ncols + 31) / 32 : 1);
    DevBuf<int32_t> cnt(n);
    sym_wave_kernel<false><<<(unsigned)n, 64, lds, s>>>(n, L.rp.get(), L.ci.get(), r_rp, r_ci, colmap,
                                                        drop_diag ? 1 : 0, ncols, cnt.get(), nullptr, nullptr);
    HIP_LAUNCH_CHECK();
    int64_t nnz = dev_exclusive_scan_i32(cnt.get(), C.rp.get(), n, s);
    if (nnz >= ((int64_t)1 << 31)) fail(FEM_ERR_CAPACITY, "sparse product has 2^31 or more nonzeros");
    C.nnz = nnz;
    C.ci.resize(nnz > 0 ? nnz : 1);
    sym_wave_kernel<true><<<(unsigned)n, 64, lds, s>>>(n, L.rp.get(), L.ci.get(), r_rp, r_ci, colmap,
                                                       drop_diag ? 1 : 0, ncols, cnt.get(), C.rp.get(), C.ci.get());
    HIP_LAUNCH_CHECK();
    HIP_CHECK(hipStreamSynchronize(s));
    return;
  }
  // hash table size from the average number of product terms per row: half
  // the LDS per thread (twice the blocks per CU) when rows are short; rows
  // that do not fit take the global-memory path either way
  bool small = true;
  double ravg = 0.0;  // average R row length
  if (r_rp && n > 0 && r_rows > 0) {
    int32_t rnnz = 0;
    HIP_CHECK(hipMemcpyAsync(&rnnz, r_rp + r_rows, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    ravg = (double)rnnz / (double)r_rows;
    small = ((double)L.nnz / (double)n) * ravg <= 128.0;
  }
  DevBuf<int32_t> cnt(n > 0 ? n : 1);
  // rows that fit the LDS table keep their sorted columns from the counting
  // pass in a stash of limit-sized slots (so the fill is a copy, not a second
  // hashing pass) when the stash stays within kStashBytes
  const int stride = (small ? kSymCapSmall : kSymCap) * 3 / 4;
  const bool stashing = n > 0 && (double)n * stride * sizeof(int32_t) <= kStashBytes;
  DevBuf<int32_t> stash;
  DevBuf<uint8_t> fit;
  if (stashing) {
    stash.resize((size_t)n * stride);
    fit.resize((size_t)n);
    // lanes per row from the length of what they split (the R rows, else
    // the L rows); at least 8 with 256 slots (a block's tables within 64 KB)
    const double split = r_rp ? ravg : (double)L.nnz / (double)n;
    int g = split <= 4.0 ? 4 : (split <= 8.0 ? 8 : 16);
    if (!small) g = std::max(g, 8);
    auto launch = [&](auto gc, auto capc) {
      constexpr int G = decltype(gc)::value, CAP = decltype(capc)::value;
      const unsigned nb = (unsigned)((n + kSymNT / G - 1) / (kSymNT / G));
      sym_group_stash_kernel<G, CAP><<<nb, kSymNT, 0, s>>>(n, L.rp.get(), L.ci.get(), r_rp, r_ci, colmap,
                                                           drop_diag ? 1 : 0, cnt.get(), stash.get(), fit.get());
    };
    using Small = std::integral_constant<int, kSymCapSmall>;
    using Large = std::integral_constant<int, kSymCap>;
    switch (g) {
      case 4: launch(std::integral_constant<int, 4>{}, Small{}); break;  // g = 4 only with the small table
      case 8:
        if (small) launch(std::integral_constant<int, 8>{}, Small{});
        else launch(std::integral_constant<int, 8>{}, Large{});
        break;
      default:
        if (small) launch(std::integral_constant<int, 16>{}, Small{});
        else launch(std::integral_constant<int, 16>{}, Large{});
        break;
    }
    HIP_LAUNCH_CHECK();
  } else if (n > 0) {
    if (small)
      sym_lds_kernel<false, kSymCapSmall><<<grid_for(n, kSymBlock), kSymBlock, 0, s>>>(
          n, L.rp.get(), L.ci.get(), r_rp, r_ci, colmap, drop_diag ? 1 : 0, cnt.get(), nullptr, nullptr);
    else
      sym_lds_kernel<false, kSymCap><<<grid_for(n, kSymBlock), kSymBlock, 0, s>>>(
          n, L.rp.get(), L.ci.get(), r_rp, r_ci, colmap, drop_diag ? 1 : 0, cnt.get(), nullptr, nullptr);
    HIP_LAUNCH_CHECK();
  }
  DevBuf<int32_t> rows(n > 0 ? n : 1);
  DevBuf<int64_t> cap(n + 1);
  DevBuf<unsigned> nover_d(1);
  HIP_CHECK(hipMemsetAsync(nover_d.get(), 0, sizeof(unsigned), s));
  if (n > 0) {
    collect_overflow_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, cnt.get(), L.rp.get(), L.ci.get(), r_rp, ncols,
                                                             rows.get(), cap.get(), nover_d.get());
    HIP_LAUNCH_CHECK();
  }
  unsigned nover = 0;
  HIP_CHECK(hipMemcpyAsync(&nover, nover_d.get(), sizeof(unsigned), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  DevBuf<int64_t> off;
  DevBuf<int> scratch;
  if (nover) {
    off.resize(nover + 1);
    int64_t tot = dev_exclusive_scan_i64(cap.get(), off.get(), nover, s);
    scratch.resize(tot);
    sym_global_kernel<false><<<grid_for(nover, 64), 64, 0, s>>>(nover, rows.get(), off.get(), L.rp.get(), L.ci.get(),
                                                                r_rp, r_ci, colmap, drop_diag ? 1 : 0, scratch.get(),
                                                                cnt.get(), nullptr, nullptr);
    HIP_LAUNCH_CHECK();
  }
  int64_t nnz = dev_exclusive_scan_i32(cnt.get(), C.rp.get(), n, s);
  if (nnz >= ((int64_t)1 << 31)) fail(FEM_ERR_CAPACITY, "sparse product has 2^31 or more nonzeros");
  C.nnz = nnz;
  C.ci.resize(nnz > 0 ? nnz : 1);
  if (stashing) {
    stash_copy_kernel<<<grid_for(n * kCopyLanes, 256), 256, 0, s>>>(n, stride, fit.get(), stash.get(), C.rp.get(),
                                                                     C.ci.get());
    HIP_LAUNCH_CHECK();
  } else if (n > 0) {
    if (small)
      sym_lds_kernel<true, kSymCapSmall><<<grid_for(n, kSymBlock), kSymBlock, 0, s>>>(
          n, L.rp.get(), L.ci.get(), r_rp, r_ci, colmap, drop_diag ? 1 : 0, cnt.get(), C.rp.get(), C.ci.get());
    else
      sym_lds_kernel<true, kSymCap><<<grid_for(n, kSymBlock), kSymBlock, 0, s>>>(
          n, L.rp.get(), L.ci.get(), r_rp, r_ci, colmap, drop_diag ? 1 : 0, cnt.get(), C.rp.get(), C.ci.get());
    HIP_LAUNCH_CHECK();
  }
  if (nover) {
    sym_global_kernel<true><<<grid_for(nover, 64), 64, 0, s>>>(nover, rows.get(), off.get(), L.rp.get(), L.ci.get(),
                                                               r_rp, r_ci, colmap, drop_diag ? 1 : 0, scratch.get(),
                                                               cnt.get(), C.rp.get(), C.ci.get());
    HIP_LAUNCH_CHECK();
  }
  HIP_CHECK(hipStreamSynchronize(s));
}

void spgemm_numeric(const Csr& L, const Csr& Rm, Csr& C, hipStream_t s) {
  C.v64.resize(C.nnz > 0 ? C.nnz : 1);
  if (C.n == 0) return;
  if (use_wave_rows(C.n, C.m)) {
    num_wave_kernel<<<(unsigned)C.n, 64, 0, s>>>(C.n, L.rp.get(), L.ci.get(), L.v64.get(), Rm.rp.get(), Rm.ci.get(),
                                                  Rm.v64.get(), C.rp.get(), C.ci.get(), C.v64.get());
    HIP_LAUNCH_CHECK();
    return;
  }
  // table size: the longest row, up to 256 (longer rows sum in global
  // memory); lanes per row: the R rows' average length
  DevBuf<int> mx(1);
  HIP_CHECK(hipMemsetAsync(mx.get(), 0, sizeof(int), s));
  max_row_len_kernel<<<std::min<unsigned>(grid_for(C.n, 256), 1024u), 256, 0, s>>>(C.n, C.rp.get(), mx.get());
  HIP_LAUNCH_CHECK();
  int maxlen = 0;
  HIP_CHECK(hipMemcpyAsync(&maxlen, mx.get(), sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  const double ravg = Rm.n > 0 ? (double)Rm.nnz / (double)Rm.n : 0.0;
  const int cap = maxlen <= 32 ? 32 : (maxlen <= 64 ? 64 : (maxlen <= 128 ? 128 : 256));
  // at least cap / 16 lanes per row: a block's tables stay within 64 KB
  const int g = std::max(ravg <= 4.0 ? 4 : (ravg <= 8.0 ? 8 : 16), cap / 16);
  auto launch = [&](auto gc, auto capc) {
    constexpr int G = decltype(gc)::value, CAP = decltype(capc)::value;
    if constexpr ((size_t)(kNumNT / G) * (CAP + 1) * 12 <= 65536) {
      const unsigned nb = (unsigned)((C.n + kNumNT / G - 1) / (kNumNT / G));
      num_group_kernel<G, CAP><<<nb, kNumNT, 0, s>>>(C.n, L.rp.get(), L.ci.get(), L.v64.get(), Rm.rp.get(),
                                                     Rm.ci.get(), Rm.v64.get(), C.rp.get(), C.ci.get(), C.v64.get());
    } else {
      fail(FEM_ERR_STATE, "numeric sparse product: table size and lanes per row out of range");
    }
  };
  auto by_cap = [&](auto gc) {
    switch (cap) {
      case 32: launch(gc, std::integral_constant<int, 32>{}); break;
      case 64: launch(gc, std::integral_constant<int, 64>{}); break;
      case 128: launch(gc, std::integral_constant<int, 128>{}); break;
      default: launch(gc, std::integral_constant<int, 256>{}); break;
    }
  };
  switch (g) {
    case 4: by_cap(std::integral_constant<int, 4>{}); break;
    case 8: by_cap(std::integral_constant<int, 8>{}); break;
    default: by_cap(std::integral_constant<int, 16>{}); break;
  }
  HIP_LAUNCH_CHECK();
}

void build_sell(Csr& A, bool with_f32, hipStream_t s, int64_t sort_window, bool with_f16, bool with_f64,
                int lanes, bool split_diag) {
  if (lanes < 1 || lanes > 32 || (lanes & (lanes - 1)) != 0)
    fail(FEM_ERR_ARG, "sliced ELL: 1, 2, 4, 8, 16 or 32 lanes per row");
  A.sd64.release();
  A.sd32.release();
  A.nnz_sdiag = 0;
  A.lds_bs = 0;
  A.lds_maxc = 0;
  A.lc.release();
  A.lcs.release();
  A.lcols.release();
  if (split_diag) {
    if (A.n != A.m || lanes != 1 || sort_window > 0 || with_f16)
      fail(FEM_ERR_ARG, "sliced ELL: the diagonal is split off square matrices in row order only");
    // slices of the off-diagonal entries, then the diagonal beside them
    Csr T;
    T.n = T.m = A.n;
    DevBuf<int32_t> cnt(A.n > 0 ? A.n : 1);
    DevBuf<double> diag(A.n > 0 ? A.n : 1);
    if (A.n > 0) {
      split_diag_count_kernel<<<grid_for(A.n, 256), 256, 0, s>>>(A.n, A.rp.get(), A.ci.get(), cnt.get());
      HIP_LAUNCH_CHECK();
    }
    T.rp.resize(A.n + 1);
    T.nnz = dev_exclusive_scan_i32(cnt.get(), T.rp.get(), A.n, s);
    T.ci.resize(T.nnz > 0 ? T.nnz : 1);
    T.v64.resize(T.nnz > 0 ? T.nnz : 1);
    if (A.n > 0) {
      split_diag_fill_kernel<<<grid_for(A.n * kSplitLanes, 256), 256, 0, s>>>(
          A.n, A.rp.get(), A.ci.get(), A.v64.get(), T.rp.get(), T.ci.get(), T.v64.get(), diag.get());
      HIP_LAUNCH_CHECK();
    }
    build_sell(T, with_f32, s, 0, false, with_f64, 1, false);
    A.nslices = T.nslices;
    A.sell_g = 1;
    A.sp = std::move(T.sp);
    A.sc = std::move(T.sc);
    A.sc16 = std::move(T.sc16);
    A.swin = std::move(T.swin);
    A.ncompact = T.ncompact;
    A.nnz_compact = T.nnz_compact;
    A.srow.release();
    A.s64 = std::move(T.s64);
    A.s32 = std::move(T.s32);
    A.sh.release();
    A.nnz_sdiag = A.nnz - T.nnz;
    if (with_f32 && A.n > 0) {
      A.sd32.resize(A.n);
      dev_f64_to_f32(diag.get(), A.sd32.get(), A.n, s);
    }
    if (with_f64) A.sd64 = std::move(diag);
    HIP_CHECK(hipStreamSynchronize(s));
    return;
  }
  const int64_t rows = 64 / lanes;
  const int64_t ns = (A.n + rows - 1) / rows;
  A.nslices = ns;
  A.sell_g = lanes;
  A.sp.resize(ns + 1);
  A.srow.release();
  if (sort_window > 0 && A.n > 0) {
    // rows reordered, longest first, inside windows of sort_window rows: the
    // slices of rows of mixed lengths (prolongators) then carry little padding
    if ((A.n - 1) / sort_window >= ((int64_t)1 << 24)) fail(FEM_ERR_CAPACITY, "too many rows to sort by length");
    DevBuf<uint32_t> key(A.n), key2(A.n);
    DevBuf<int32_t> idx(A.n);
    A.srow.resize(A.n);
    sell_sort_keys_kernel<<<grid_for(A.n, 256), 256, 0, s>>>(A.n, A.rp.get(), sort_window, key.get(), idx.get());
    HIP_LAUNCH_CHECK();
    dev_sort_pairs_u32_i32(key.get(), key2.get(), idx.get(), A.srow.get(), (size_t)A.n, 32, s);
  }
  const int32_t* srow = A.srow.size() ? A.srow.get() : nullptr;
  DevBuf<int32_t> w(ns > 0 ? ns : 1);
  if (ns > 0) {
    sell_width_kernel<<<grid_for(ns, 256), 256, 0, s>>>(A.n, ns, lanes, A.rp.get(), srow, w.get());
    HIP_LAUNCH_CHECK();
  }
  int64_t tot = dev_exclusive_scan_i32(w.get(), A.sp.get(), ns, s);
  if (tot >= ((int64_t)1 << 31)) fail(FEM_ERR_CAPACITY, "sliced-ELL copy has 2^31 or more entries");
  A.sc.resize(tot > 0 ? tot : 1);
  A.sc16.resize(tot > 0 ? tot : 1);
  A.swin.resize(ns > 0 ? ns * kSellWindows : kSellWindows);
  if (with_f64) A.s64.resize(tot > 0 ? tot : 1); else A.s64.release();
  if (with_f32) A.s32.resize(tot > 0 ? tot : 1); else A.s32.release();
  if (with_f16) A.sh.resize(tot > 0 ? tot : 1); else A.sh.release();
  A.ncompact = 0;
  A.nnz_compact = 0;
  if (ns > 0) {
    DevBuf<int32_t> cflag(ns), cscan(ns + 1);
    sell_window_range_kernel<<<grid_for(ns * 64, 256), 256, 0, s>>>(A.n, ns, lanes, A.rp.get(), A.ci.get(), srow,
                                                                    A.swin.get(), cflag.get());
    HIP_LAUNCH_CHECK();
    sell_window_kernel<<<grid_for(ns * 64, 256), 256, 0, s>>>(A.n, ns, lanes, A.rp.get(), A.ci.get(), srow, A.swin.get(),
                                                         cflag.get());
    HIP_LAUNCH_CHECK();
    A.ncompact = dev_exclusive_scan_i32(cflag.get(), cscan.get(), ns, s);
    DevBuf<int64_t> snnz(ns), sscan(ns + 1);
    sell_compact_nnz_kernel<<<grid_for(ns, 256), 256, 0, s>>>(A.n, ns, lanes, A.rp.get(), srow, cflag.get(),
                                                              snnz.get());
    HIP_LAUNCH_CHECK();
    A.nnz_compact = dev_exclusive_scan_i64(snnz.get(), sscan.get(), ns, s);
    // one lane per row in row order: slices of short rows fill through LDS
    const bool staged = lanes == 1 && srow == nullptr;
    if (staged) {
      sell_fill_lds_kernel<<<(unsigned)ns, 64, 0, s>>>(A.n, ns, A.rp.get(), A.ci.get(), A.v64.get(), A.sp.get(),
                                                       A.swin.get(), A.sc.get(), A.sc16.get(),
                                                       with_f64 ? A.s64.get() : nullptr,
                                                       with_f32 ? A.s32.get() : nullptr,
                                                       with_f16 ? A.sh.get() : nullptr);
      HIP_LAUNCH_CHECK();
    }
    sell_fill_kernel<<<grid_for(ns * 64, 256), 256, 0, s>>>(A.n, ns, lanes, A.rp.get(), A.ci.get(), A.v64.get(),
                                                            srow, A.sp.get(), A.swin.get(), A.sc.get(), A.sc16.get(),
                                                            with_f64 ? A.s64.get() : nullptr,
                                                            with_f32 ? A.s32.get() : nullptr,
                                                            with_f16 ? A.sh.get() : nullptr, staged ? 1 : 0);
    HIP_LAUNCH_CHECK();
  }
  HIP_CHECK(hipStreamSynchronize(s));
}

void build_lds(Csr& A, int bs, hipStream_t s) {
  A.lds_bs = 0;
  A.lds_maxc = 0;
  A.lc.release();
  A.lcs.release();
  A.lcols.release();
  if (A.nslices <= 0 || A.sell_g != 1 || bs <= 0) return;
  const int64_t ns = A.nslices;
  const int64_t nb = (ns + bs - 1) / bs;
  DevBuf<int32_t> lo(nb), nw(nb), cnt(nb), wide(nb);
  DevBuf<int> flag(4);  // max bitmap words, max distinct columns, wide blocks, sort overflow
  dev_fill_i32(flag.get(), 0, 4, s);
  lds_range_kernel<<<(unsigned)nb, kLdsThreads, 0, s>>>(ns, bs, A.sp.get(), A.sc.get(), lo.get(), nw.get(),
                                                        flag.get(), wide.get());
  HIP_LAUNCH_CHECK();
  int hf[4] = {0, 0, 0, 0};
  HIP_CHECK(hipMemcpyAsync(hf, flag.get(), sizeof(hf), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  const int nwide = hf[2];
  const size_t lds = 2 * sizeof(uint32_t) * (size_t)(hf[0] > 0 ? hf[0] : 1);
  const size_t slds = sizeof(int32_t) * (kLdsHash + kLdsMaxCols);
  lds_cols_kernel<false><<<(unsigned)nb, kLdsThreads, lds, s>>>(ns, bs, A.sp.get(), A.sc.get(), lo.get(), nw.get(),
                                                               cnt.get(), nullptr, nullptr, nullptr);
  HIP_LAUNCH_CHECK();
  if (nwide > 0) {
    lds_hash_kernel<false><<<(unsigned)nwide, kLdsThreads, slds, s>>>(ns, bs, wide.get(), A.sp.get(), A.sc.get(),
                                                                      cnt.get(), nullptr, nullptr, nullptr,
                                                                      flag.get() + 3);
    HIP_LAUNCH_CHECK();
  }
  max_i32_kernel<<<grid_for(nb, 256), 256, 0, s>>>(nb, cnt.get(), flag.get() + 1);
  HIP_LAUNCH_CHECK();
  HIP_CHECK(hipMemcpyAsync(hf, flag.get(), sizeof(hf), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  const int maxc = hf[1];
  if (hf[3] || maxc > kLdsMaxCols) return;  // a block too large to stage
  DevBuf<int32_t> lcs(nb + 1);
  const int64_t tot = dev_exclusive_scan_i32(cnt.get(), lcs.get(), nb, s);
  DevBuf<int32_t> lcols(tot > 0 ? tot : 1);
  const int64_t ent = (int64_t)A.sc.size();
  DevBuf<uint16_t> lc(ent > 0 ? ent : 1);
  lds_cols_kernel<true><<<(unsigned)nb, kLdsThreads, lds, s>>>(ns, bs, A.sp.get(), A.sc.get(), lo.get(), nw.get(),
                                                              nullptr, lcs.get(), lcols.get(), lc.get());
  HIP_LAUNCH_CHECK();
  if (nwide > 0) {
    lds_hash_kernel<true><<<(unsigned)nwide, kLdsThreads, slds, s>>>(ns, bs, wide.get(), A.sp.get(), A.sc.get(),
                                                                     nullptr, lcs.get(), lcols.get(), lc.get(),
                                                                     flag.get() + 3);
    HIP_LAUNCH_CHECK();
  }
  HIP_CHECK(hipStreamSynchronize(s));
  A.lc = std::move(lc);
  A.lcs = std::move(lcs);
  A.lcols = std::move(lcols);
  A.lds_maxc = maxc;
  A.lds_bs = bs;
}

void csr_transpose(const Csr& P, Csr& R, hipStream_t s) {
  const int64_t nnz = P.nnz;
  R.n = P.m;
  R.m = P.n;
  R.nnz = nnz;
  R.rp.resize(R.n + 1);
  R.ci.resize(nnz > 0 ? nnz : 1);
  R.v64.resize(nnz > 0 ? nnz : 1);
  DevBuf<uint64_t> key(nnz > 0 ? nnz : 1), key2(nnz > 0 ? nnz : 1);
  DevBuf<int32_t> idx(nnz > 0 ? nnz : 1), idx2(nnz > 0 ? nnz : 1);
  if (P.n > 0) {
    transpose_keys_kernel<<<grid_for(P.n, 256), 256, 0, s>>>(P.n, P.rp.get(), P.ci.get(), key.get(), idx.get());
    HIP_LAUNCH_CHECK();
  }
  dev_sort_pairs_u64_i32(key.get(), key2.get(), idx.get(), idx2.get(), (size_t)nnz, 32 + bits_for(P.m), s);
  if (nnz > 0) {
    transpose_fill_kernel<<<grid_for(nnz, 256), 256, 0, s>>>(nnz, key2.get(), idx2.get(), P.v64.get(), R.ci.get(),
                                                             R.v64.get());
    HIP_LAUNCH_CHECK();
  }
  transpose_ptr_kernel<<<grid_for(R.n + 1, 256), 256, 0, s>>>(R.n, key2.get(), nnz, R.rp.get());
  HIP_LAUNCH_CHECK();
  HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace fem
