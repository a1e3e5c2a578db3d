// Smoothed-aggregation AMG setup on the device.
//
// Reference algorithm (src/core/cuda/amg.cu:79-144, SmoothedMG_AMG_Level::
// createNextLevel src/core/cuda/smoothedMG_amg_level.cu:403-494):
//   per level: randomized distance-2 MIS (randomizedMIS_GPU.cu:154-272) ->
//   node-to-aggregate allocation with removal of aggregates smaller than 9
//   (misHelpers.cu:13-75,513-601,744) -> tentative T(i, agg(i)) = 1 ->
//   P = (I - omega D^-1 A) T (smoothedMG_amg_level.cu:321-387) -> R = P^T ->
//   A_c = R (A P) (:390-400); the next level aggregates the induced aggregate
//   graph (misHelpers.cu:1078); the coarsest level (< topSize rows) is solved
//   directly (amg.cu:101-107).
// The MIS kernels below keep the reference's arithmetic (Tausworthe steps,
// value % 1000000, 1000001 for members, ties to the larger origin) with the
// time() seed replaced by a SplitMix64 stream, so the aggregation is a
// deterministic function of (graph, seed) and matches the oracle bit for bit.
#include <algorithm>
#include <climits>
#include <chrono>
#include <cmath>
#include <vector>

#include "fem_internal.hpp"
#include "half.hpp"
#include "kernels.hpp"

namespace fem {

namespace {

constexpr int kMisThreads = 32768;  // Generate_Randoms_Kernel: 128 x 256 generator threads
constexpr int kMisInMis = 1000001;
// largest coarsest level inverted densely on the host (5n^3/6 multiply-adds)
constexpr int64_t kMaxCoarse = 2048;

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// seed of generator thread i (replaces rand() after srand(time(NULL)))
inline uint32_t mis_seed(uint64_t seed, uint32_t i) {
  return (uint32_t)(mix64(seed * 0x9E3779B97F4A7C15ull + (uint64_t)(i + 1) * 0xD1B54A32D192ED03ull) >> 32);
}

__device__ __forceinline__ uint32_t taus_step(uint32_t z) {
  uint32_t b = (((z << 13) ^ z) >> 19);
  return ((z & 0xffffffffu) << 12) ^ b;
}

__global__ void mis_randoms_kernel(int64_t size, int iterations, uint32_t* __restrict__ randoms,
                                   const uint32_t* __restrict__ seeds) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= kMisThreads) return;
  uint32_t z = seeds[idx];
  int64_t offset = idx;
  for (int i = 0; i < iterations; ++i) {
    if (offset < size) {
      z = taus_step(z);
      randoms[offset] = z;
      offset += kMisThreads;
    }
  }
}

__global__ void mis_first_init_kernel(int64_t size, const uint32_t* __restrict__ randoms, int* __restrict__ best,
                                      int* __restrict__ origin) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= size) return;
  origin[idx] = (int)idx;
  best[idx] = (int)(randoms[idx] % 1000000u);
}

__global__ void mis_init_kernel(int64_t size, uint32_t* __restrict__ randoms, int* __restrict__ best,
                                int* __restrict__ origin, const int* __restrict__ mis) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= size) return;
  uint32_t z = taus_step(randoms[idx]);
  origin[idx] = (int)idx;
  int status = mis[idx];
  best[idx] = (status == -1) ? (int)(z % 1000000u) : (status == 1 ? kMisInMis : 0);
  randoms[idx] = z;
}

// one hop of (best value, origin) max propagation; final hop also decides
template <bool kFinal>
__global__ void mis_iterate_kernel(int64_t size, const int* __restrict__ origin_in, int* __restrict__ origin_out,
                                   const int* __restrict__ best_in, int* __restrict__ best_out,
                                   const int32_t* __restrict__ adj_ptr, const int32_t* __restrict__ adj,
                                   int* __restrict__ mis, int* __restrict__ incomplete) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= size) return;
  int best = best_in[idx];
  int origin = origin_in[idx];
  if (best < kMisInMis) {
    for (int i = adj_ptr[idx]; i < adj_ptr[idx + 1]; ++i) {
      int nb = adj[i];
      int ch = best_in[nb];
      int cho = origin_in[nb];
      if (ch > 0 && ch == best && cho > origin) origin = cho;
      if (ch > best) {
        best = ch;
        origin = cho;
      }
    }
  }
  if (!kFinal) {
    best_out[idx] = best;
    origin_out[idx] = origin;
  } else {
    int st = -1;
    if (origin == (int)idx) st = 1;
    else if (best == kMisInMis) st = 0;
    mis[idx] = st;
    if (st == -1) *incomplete = 1;
  }
}

// partIn[i] = aggregate id of a root (rank of the root in index order), -1 else
__global__ void roots_kernel(int64_t n, const int* __restrict__ mis, const int32_t* __restrict__ scan,
                             int* __restrict__ part, int* __restrict__ aggregated) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  bool root = mis[i] == 1;
  part[i] = root ? scan[i] : -1;  // exclusive scan == inclusive - 1 for roots
  aggregated[i] = root ? 1 : 0;
}

// rule 0: aggregate of the first aggregated neighbour in adjacency order
//         (what the reference's allocateNodesKernel computes);
// rule 1: most frequent neighbouring aggregate, ties to the smallest id.
// Rule 1 counts over the row's neighbour aggregates once per distinct one;
// rows of at most kAllocCap neighbours read them from lane-interleaved LDS
// (entry t of lane j at t * kAllocNT + j), longer rows from memory.
constexpr int kAllocNT = 64;
constexpr int kAllocCap = 32;
__global__ void __launch_bounds__(kAllocNT) allocate_kernel(int64_t n, int rule, const int32_t* __restrict__ adj_ptr,
                                                           const int32_t* __restrict__ adj,
                                                           const int* __restrict__ part_in,
                                                           int* __restrict__ part_out,
                                                           int* __restrict__ aggregated) {
  __shared__ int s_nb[kAllocCap * kAllocNT];
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (aggregated[i] != 0) {
    part_out[i] = part_in[i];
    return;
  }
  int add_to = -1;
  const int p0 = adj_ptr[i], p1 = adj_ptr[i + 1];
  if (rule == 0) {
    for (int p = p0; p < p1; ++p) {
      int c = part_in[adj[p]];
      if (c != -1) {
        add_to = c;
        break;
      }
    }
  } else {
    // the row's neighbour aggregates, in adjacency order
    auto count_best = [&](auto nb) {
      const int deg = p1 - p0;
      int best_cnt = 0;
      for (int p = 0; p < deg; ++p) {
        const int c = nb(p);
        if (c == -1) continue;
        bool seen = false;
        for (int q = 0; q < p; ++q)
          if (nb(q) == c) { seen = true; break; }
        if (seen) continue;
        int cnt = 0;
        for (int q = p; q < deg; ++q) cnt += (nb(q) == c);
        if (cnt > best_cnt || (cnt == best_cnt && c < add_to)) {
          best_cnt = cnt;
          add_to = c;
        }
      }
    };
    if (p1 - p0 <= kAllocCap) {
      int* const lb = s_nb + threadIdx.x;
      for (int p = p0; p < p1; ++p) lb[(p - p0) * kAllocNT] = part_in[adj[p]];
      count_best([lb](int t) { return lb[t * kAllocNT]; });
    } else {
      count_best([=](int t) { return part_in[adj[p0 + t]]; });
    }
  }
  part_out[i] = add_to;
  if (add_to != -1) aggregated[i] = 1;
}

// Reductions to one device word: a capped grid strides over the array and
// each block adds one atomic (atomics on one address serialise in L2, about
// 10 ns each, so one per wavefront of a 4M-node pass costs 0.7 ms)
constexpr int kReduceBlocks = 1024;

__global__ void __launch_bounds__(256) count_unallocated_kernel(int64_t n, const int* __restrict__ aggregated,
                                                                unsigned* __restrict__ cnt) {
  __shared__ unsigned s_part[4];
  unsigned k = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    k += aggregated[i] == 0 ? 1u : 0u;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) k += __shfl_xor(k, off, 64);
  if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = k;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned t = s_part[0] + s_part[1] + s_part[2] + s_part[3];
    if (t) atomicAdd(cnt, t);
  }
}

__global__ void part_sizes_kernel(int64_t n, const int* __restrict__ part, int* __restrict__ sizes) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && part[i] >= 0) atomicAdd(&sizes[part[i]], 1);
}

__global__ void runt_stencil_kernel(int64_t np, const int* __restrict__ sizes, int min_size, int32_t* __restrict__ st) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < np) st[p] = sizes[p] < min_size ? 1 : 0;
}

// removeRuntyPartsKernel: drop runts, renumber survivors densely
__global__ void runt_apply_kernel(int64_t n, int* __restrict__ part, const int32_t* __restrict__ st,
                                  const int32_t* __restrict__ excl, int* __restrict__ aggregated) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int p = part[i];
  if (st[p]) {
    part[i] = -1;
    aggregated[i] = 0;
  } else {
    part[i] = p - excl[p];  // excl = exclusive count of removed ids below p
    aggregated[i] = 1;
  }
}

__global__ void stuck_flag_kernel(int64_t n, const int* __restrict__ aggregated, int32_t* __restrict__ flag) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flag[i] = aggregated[i] == 0 ? 1 : 0;
}

__global__ void stuck_assign_kernel(int64_t n, int base, const int32_t* __restrict__ flag,
                                    const int32_t* __restrict__ excl, int* __restrict__ part) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && flag[i]) part[i] = base + excl[i];
}

__global__ void __launch_bounds__(256) max_kernel(int64_t n, const int* __restrict__ part, int* __restrict__ mx) {
  __shared__ int s_part[4];
  int v = INT_MIN;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    v = max(v, part[i]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
  if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) atomicMax(mx, max(max(s_part[0], s_part[1]), max(s_part[2], s_part[3])));
}

unsigned reduce_grid(int64_t n) { return std::min<unsigned>(grid_for(n, 256), (unsigned)kReduceBlocks); }

unsigned read_counter(DevBuf<unsigned>& c, hipStream_t s) {
  unsigned v = 0;
  HIP_CHECK(hipMemcpyAsync(&v, c.get(), sizeof(unsigned), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  return v;
}

void randomized_mis(int64_t n, const int32_t* adj_ptr, const int32_t* adj, int depth, unsigned long long seed,
                    int* mis, hipStream_t s) {
  std::vector<uint32_t> seeds_h(kMisThreads);
  for (int i = 0; i < kMisThreads; ++i) seeds_h[i] = mis_seed(seed, (uint32_t)i);
  DevBuf<uint32_t> seeds;
  seeds.upload(seeds_h.data(), kMisThreads, s);
  DevBuf<uint32_t> randoms(n);
  DevBuf<int> best_a(n), best_b(n), orig_a(n), orig_b(n);
  DevBuf<int> incomplete(1);
  dev_fill_i32(mis, -1, n, s);
  int iterations = (int)((n + kMisThreads - 1) / kMisThreads);
  mis_randoms_kernel<<<kMisThreads / 256, 256, 0, s>>>(n, iterations, randoms.get(), seeds.get());
  HIP_LAUNCH_CHECK();
  mis_first_init_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, randoms.get(), best_a.get(), orig_a.get());
  HIP_LAUNCH_CHECK();
  bool first = true;
  for (int round = 0;; ++round) {
    if (round > 100000) fail(FEM_ERR_NUMERIC, "randomized MIS did not complete");
    if (!first) {
      mis_init_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, randoms.get(), best_a.get(), orig_a.get(), mis);
      HIP_LAUNCH_CHECK();
    }
    first = false;
    HIP_CHECK(hipMemsetAsync(incomplete.get(), 0, sizeof(int), s));
    int* bi = best_a.get(); int* bo = best_b.get();
    int* oi = orig_a.get(); int* oo = orig_b.get();
    for (int i = 0; i < depth; ++i) {
      if (i < depth - 1) {
        mis_iterate_kernel<false><<<grid_for(n, 256), 256, 0, s>>>(n, oi, oo, bi, bo, adj_ptr, adj, mis,
                                                                  incomplete.get());
        std::swap(bi, bo);
        std::swap(oi, oo);
      } else {
        mis_iterate_kernel<true><<<grid_for(n, 256), 256, 0, s>>>(n, oi, oo, bi, bo, adj_ptr, adj, mis,
                                                                 incomplete.get());
      }
      HIP_LAUNCH_CHECK();
    }
    // Initialize_Kernel rewrites best/origin completely each round, so which
    // buffer pair is "in" at the start of a round does not change the result.
    int inc = 0;
    HIP_CHECK(hipMemcpyAsync(&inc, incomplete.get(), sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (!inc) break;
  }
}

// ---- prolongator, diagonal, power iteration ---------------------------------------

__global__ void diag_kernel(int64_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                            const double* __restrict__ v, double* __restrict__ dinv, int* __restrict__ bad) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double d = 0.0;
  int lo = rp[i], hi = rp[i + 1];
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (ci[mid] < (int)i) lo = mid + 1; else hi = mid;
  }
  if (lo < rp[i + 1] && ci[lo] == (int)i) d = v[lo];
  if (!(d > 0.0)) {
    atomicAdd(bad, 1);
    dinv[i] = 0.0;
  } else {
    dinv[i] = 1.0 / d;
  }
}

// P(i, J) = delta(J, agg(i)) - omega / d_i * sum_{k: agg(k) = J} a_ik, summed in
// ascending k then the tentative 1 (order of the reference's reduce_by_key).
// Rows of at most kProRowCap entries accumulate in LDS, lane-interleaved
// (entry t of lane j at t * kProNT + j), with the columns searched there;
// longer rows accumulate in place.  Either way each entry takes its terms in
// the order of the row of A, then the tentative 1.
constexpr int kProNT = 64;
constexpr int kProRowCap = 32;
__global__ void __launch_bounds__(kProNT) prolongator_kernel(int64_t n, const int32_t* __restrict__ a_rp,
                                                           const int32_t* __restrict__ a_ci,
                                                           const double* __restrict__ a_v,
                                                           const double* __restrict__ dinv,
                                                           const int32_t* __restrict__ agg, double omega,
                                                           const int32_t* __restrict__ p_rp,
                                                           const int32_t* __restrict__ p_ci, double* __restrict__ p_v) {
  __shared__ int s_c[kProRowCap * kProNT];
  __shared__ double s_v[kProRowCap * kProNT];
  int64_t i = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;  // XCD-contiguous rows
  if (i >= n) return;
  const int o = p_rp[i], len = p_rp[i + 1] - o;
  const double d = (dinv[i] != 0.0) ? 1.0 / dinv[i] : 0.0;
  auto row = [&](auto cols, auto vals) {
    auto slot = [&](int J) {
      int lo = 0, hi = len;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (cols(mid) < J) lo = mid + 1; else hi = mid;
      }
      return lo;
    };
    for (int q = a_rp[i]; q < a_rp[i + 1]; ++q) vals(slot(agg[a_ci[q]])) += (-omega * a_v[q]) / d;
    vals(slot(agg[i])) += 1.0;
  };
  if (len <= kProRowCap) {
    int* const lc = s_c + threadIdx.x;
    double* const lv = s_v + threadIdx.x;
    for (int t = 0; t < len; ++t) {
      lc[t * kProNT] = p_ci[o + t];
      lv[t * kProNT] = 0.0;
    }
    row([lc](int t) -> int& { return lc[t * kProNT]; }, [lv](int t) -> double& { return lv[t * kProNT]; });
    for (int t = 0; t < len; ++t) p_v[o + t] = lv[t * kProNT];
  } else {
    for (int t = 0; t < len; ++t) p_v[o + t] = 0.0;
    const int32_t* gc = p_ci + o;
    double* gv = p_v + o;
    row([gc](int t) -> const int32_t& { return gc[t]; }, [gv](int t) -> double& { return gv[t]; });
  }
}

// per-block partial sums of v_i^2; with dinv, v is first scaled in place
// (v_i *= dinv_i) and the sum taken over the scaled values
__global__ void sumsq_kernel(int64_t n, double* __restrict__ v, const double* __restrict__ dinv,
                             double* __restrict__ part) {
  __shared__ double s_buf[kNT];
  double a = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kNT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kNT) {
    double x = v[i];
    if (dinv) {
      x *= dinv[i];
      v[i] = x;
    }
    a += x * x;
  }
  double t = block_sum(a, s_buf);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}


__global__ void hash_vector_kernel(int64_t n, unsigned long long seed, double* __restrict__ v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t z = (uint64_t)seed + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  v[i] = 0.5 + (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

// one block: nrm = sqrt(sum of the nb partials, summed in index order)
__global__ void norm_finish_kernel(int nb, const double* __restrict__ part, double* __restrict__ nrm) {
  if (threadIdx.x != 0) return;
  double t = 0.0;
  // in index order; the loads of a chunk are issued together
  int b = 0;
  for (; b + 16 <= nb; b += 16) {
    double c[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) c[j] = part[b + j];
#pragma unroll
    for (int j = 0; j < 16; ++j) t += c[j];
  }
  for (; b < nb; ++b) t += part[b];
  nrm[0] = sqrt(t);
}

// v *= 1 / nrm when nrm > 0 (a zero or NaN norm leaves v as it is)
__global__ void scale_by_norm_kernel(int64_t n, const double* __restrict__ nrm, double* __restrict__ v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const double a = nrm[0];
  if (i < n && a > 0) v[i] *= 1.0 / a;
}

// Spectral radius of D^-1 A by power iteration (upper part of the spectrum),
// on the device with one read-back at the end.
double estimate_rho(const Csr& A, const double* dinv, hipStream_t s, int iters = 15) {
  const int64_t n = A.n;
  const int nb = 256;
  DevBuf<double> v(n), w(n), part(nb), nrm(iters + 1);
  hash_vector_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, 12345ull, v.get());
  HIP_LAUNCH_CHECK();
  auto norm_into = [&](double* x, const double* scale, double* out) {
    sumsq_kernel<<<nb, kNT, 0, s>>>(n, x, scale, part.get());
    HIP_LAUNCH_CHECK();
    norm_finish_kernel<<<1, 64, 0, s>>>(nb, part.get(), out);
    HIP_LAUNCH_CHECK();
  };
  norm_into(v.get(), nullptr, nrm.get());
  scale_by_norm_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, nrm.get(), v.get());
  HIP_LAUNCH_CHECK();
  for (int it = 0; it < iters; ++it) {
    spmv64(A, v.get(), w.get(), s);
    norm_into(w.get(), dinv, nrm.get() + 1 + it);  // w = D^-1 A v and its norm
    scale_by_norm_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, nrm.get() + 1 + it, w.get());
    HIP_LAUNCH_CHECK();
    std::swap(v, w);
  }
  std::vector<double> h = nrm.to_host(s);
  // the last norm, or the first that was not positive (the iteration stops there)
  double rho = 1.0;
  for (int it = 0; it < iters; ++it) {
    rho = h[1 + it];
    if (!(rho > 0)) break;
  }
  return rho;
}

__global__ void lower_bound_u32_kernel(int64_t m, const uint32_t* __restrict__ keys, int64_t nk,
                                       int32_t* __restrict__ ptr) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v > m) return;
  int64_t lo = 0, hi = nk;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if ((int64_t)keys[mid] < v) lo = mid + 1; else hi = mid;
  }
  ptr[v] = (int32_t)lo;
}

void lower_bound_u32(int64_t m, const uint32_t* keys, int64_t nk, int32_t* ptr, hipStream_t s) {
  lower_bound_u32_kernel<<<grid_for(m + 1, 256), 256, 0, s>>>(m, keys, nk, ptr);
  HIP_LAUNCH_CHECK();
}

// members of each aggregate, ascending node id: CSR (nagg rows)
void aggregate_members(const int32_t* agg, int64_t n, int64_t nagg, Csr& M, hipStream_t s) {
  DevBuf<uint32_t> key(n), key2(n);
  DevBuf<int32_t> idx(n);
  HIP_CHECK(hipMemcpyAsync(key.get(), agg, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, s));
  dev_iota_i32(idx.get(), n, s);
  M.n = nagg;
  M.m = n;
  M.nnz = n;
  M.ci.resize(n);
  dev_sort_pairs_u32_i32(key.get(), key2.get(), idx.get(), M.ci.get(), n, 32, s);
  M.rp.resize(nagg + 1);
  lower_bound_u32(nagg, key2.get(), n, M.rp.get(), s);
}

// Dense inverse of a small SPD matrix (Cholesky; LU with partial pivoting if
// Cholesky breaks down).  a is row-major n x n, replaced by its inverse.
bool dense_inverse(std::vector<double>& a, int64_t n) {
  std::vector<double> L(a);
  // Cholesky reads the lower triangle only: take it for symmetric matrices
  // (Galerkin products are symmetric to rounding), LU for the rest
  double amax = 0.0;
  for (int64_t i = 0; i < n * n; ++i) amax = std::max(amax, std::fabs(a[i]));
  bool spd = true;
  for (int64_t i = 0; i < n && spd; ++i)
    for (int64_t j = 0; j < i; ++j)
      if (std::fabs(a[i * n + j] - a[j * n + i]) > 1e-12 * amax) { spd = false; break; }
  for (int64_t j = 0; j < n && spd; ++j) {
    double d = L[j * n + j];
    for (int64_t k = 0; k < j; ++k) d -= L[j * n + k] * L[j * n + k];
    if (!(d > 0.0)) { spd = false; break; }
    d = std::sqrt(d);
    L[j * n + j] = d;
    for (int64_t i = j + 1; i < n; ++i) {
      double v = L[i * n + j];
      for (int64_t k = 0; k < j; ++k) v -= L[i * n + k] * L[j * n + k];
      L[i * n + j] = v / d;
    }
  }
  std::vector<double> inv(n * n, 0.0);
  if (spd) {
    // W = L^-1 row by row, then A^-1 = W^T W; every inner loop runs along a
    // row (unit stride), about 5n^3/6 multiply-adds in all
    std::vector<double> W(n * n, 0.0), r(n);
    for (int64_t i = 0; i < n; ++i) {
      std::fill(r.begin(), r.begin() + i, 0.0);
      for (int64_t k = 0; k < i; ++k) {
        const double lik = L[i * n + k];
        const double* wk = &W[k * n];
        for (int64_t j = 0; j <= k; ++j) r[j] += lik * wk[j];
      }
      const double di = 1.0 / L[i * n + i];
      for (int64_t j = 0; j < i; ++j) W[i * n + j] = -r[j] * di;
      W[i * n + i] = di;
    }
    for (int64_t k = 0; k < n; ++k) {
      const double* wk = &W[k * n];
      for (int64_t i = 0; i <= k; ++i) {
        const double wki = wk[i];
        double* row = &inv[i * n];
        for (int64_t j = 0; j <= k; ++j) row[j] += wki * wk[j];
      }
    }
    a.swap(inv);
    return true;
  }
  // LU with partial pivoting
  std::vector<double> U(a);
  std::vector<int64_t> piv(n);
  for (int64_t i = 0; i < n; ++i) piv[i] = i;
  for (int64_t j = 0; j < n; ++j) {
    int64_t p = j;
    for (int64_t i = j + 1; i < n; ++i)
      if (std::fabs(U[i * n + j]) > std::fabs(U[p * n + j])) p = i;
    if (U[p * n + j] == 0.0) return false;
    if (p != j) {
      for (int64_t k = 0; k < n; ++k) std::swap(U[p * n + k], U[j * n + k]);
      std::swap(piv[p], piv[j]);
    }
    for (int64_t i = j + 1; i < n; ++i) {
      double f = U[i * n + j] / U[j * n + j];
      U[i * n + j] = f;
      for (int64_t k = j + 1; k < n; ++k) U[i * n + k] -= f * U[j * n + k];
    }
  }
  std::vector<double> y(n);
  for (int64_t c = 0; c < n; ++c) {
    for (int64_t i = 0; i < n; ++i) {
      double v = (piv[i] == c) ? 1.0 : 0.0;
      for (int64_t k = 0; k < i; ++k) v -= U[i * n + k] * y[k];
      y[i] = v;
    }
    for (int64_t i = n - 1; i >= 0; --i) {
      double v = y[i];
      for (int64_t k = i + 1; k < n; ++k) v -= U[i * n + k] * inv[k * n + c];
      inv[i * n + c] = v / U[i * n + i];
    }
  }
  a.swap(inv);
  return true;
}

// Pattern of A without the diagonal (the aggregation graph on the full pattern).
template <bool kFill>
__global__ void offdiag_kernel(int64_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                               int32_t* __restrict__ cnt, const int32_t* __restrict__ gp, int32_t* __restrict__ gj) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int c = 0;
  const int o = kFill ? gp[i] : 0;
  for (int q = rp[i]; q < rp[i + 1]; ++q) {
    const int j = ci[q];
    if (j == (int)i) continue;
    if (kFill) gj[o + c] = j;
    ++c;
  }
  if (!kFill) cnt[i] = c;
}

void offdiag_graph(const Csr& A, Csr& G, hipStream_t s) {
  const int64_t n = A.n;
  DevBuf<int32_t> cnt(n > 0 ? n : 1);
  if (n > 0) {
    offdiag_kernel<false><<<grid_for(n, 256), 256, 0, s>>>(n, A.rp.get(), A.ci.get(), cnt.get(), nullptr, nullptr);
    HIP_LAUNCH_CHECK();
  }
  G.n = G.m = n;
  G.rp.resize(n + 1);
  G.nnz = dev_exclusive_scan_i32(cnt.get(), G.rp.get(), n, s);
  G.ci.resize(G.nnz > 0 ? G.nnz : 1);
  if (n > 0) {
    offdiag_kernel<true><<<grid_for(n, 256), 256, 0, s>>>(n, A.rp.get(), A.ci.get(), nullptr, G.rp.get(), G.ci.get());
    HIP_LAUNCH_CHECK();
  }
}

// Prolongator pattern for rows of at most kAggRowCap entries: row i holds the
// sorted distinct aggregates agg(j) of the columns j of row i of A (what the
// sparse product with the aggregate map gives), by insertion into a sorted
// list in LDS.  Longer rows are flagged in *over.
constexpr int kAggRowCap = 32;
constexpr int kAggNT = 128;

template <bool kFill>
__global__ void __launch_bounds__(kAggNT) agg_pattern_kernel(int64_t n, const int32_t* __restrict__ rp,
                                                            const int32_t* __restrict__ ci,
                                                            const int32_t* __restrict__ agg, int32_t* __restrict__ cnt,
                                                            const int32_t* __restrict__ pp, int32_t* __restrict__ pj,
                                                            int* __restrict__ over) {
  __shared__ int32_t buf[kAggNT * kAggRowCap];
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // lane-interleaved lists (entry t of lane j at t * kAggNT + j): the lanes'
  // scans and shifts address distinct banks
  int32_t* const lb = buf + threadIdx.x;
  auto lst = [lb](int t) -> int32_t& { return lb[t * kAggNT]; };
  const int s0 = rp[i], e0 = rp[i + 1];
  if (e0 - s0 > kAggRowCap) {
    if (!kFill) {
      cnt[i] = 0;
      over[0] = 1;
    }
    return;
  }
  int c = 0;
  for (int q = s0; q < e0; ++q) {
    const int a = agg[ci[q]];
    int pos = 0;
    while (pos < c && lst(pos) < a) ++pos;
    if (pos < c && lst(pos) == a) continue;
    for (int t = c; t > pos; --t) lst(t) = lst(t - 1);
    lst(pos) = a;
    ++c;
  }
  if (kFill) {
    const int o = pp[i];
    for (int t = 0; t < c; ++t) pj[o + t] = lst(t);
  } else {
    cnt[i] = c;
  }
}

// pattern of the prolongator (rows of A mapped through agg, sorted distinct);
// false when a row of A is longer than kAggRowCap (the caller then uses the
// general symbolic product)
bool agg_pattern(const Csr& A, const int32_t* agg, int64_t nagg, Csr& P, hipStream_t s) {
  const int64_t n = A.n;
  DevBuf<int32_t> cnt(n > 0 ? n : 1);
  DevBuf<int> over(1);
  dev_fill_i32(over.get(), 0, 1, s);
  if (n > 0) {
    agg_pattern_kernel<false><<<grid_for(n, kAggNT), kAggNT, 0, s>>>(n, A.rp.get(), A.ci.get(), agg, cnt.get(), nullptr,
                                                                    nullptr, over.get());
    HIP_LAUNCH_CHECK();
  }
  int ho = 0;
  HIP_CHECK(hipMemcpyAsync(&ho, over.get(), sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  if (ho) return false;
  P.n = n;
  P.m = nagg;
  P.rp.resize(n + 1);
  P.nnz = dev_exclusive_scan_i32(cnt.get(), P.rp.get(), n, s);
  P.ci.resize(P.nnz > 0 ? P.nnz : 1);
  if (n > 0) {
    agg_pattern_kernel<true><<<grid_for(n, kAggNT), kAggNT, 0, s>>>(n, A.rp.get(), A.ci.get(), agg, nullptr,
                                                                   P.rp.get(), P.ci.get(), nullptr);
    HIP_LAUNCH_CHECK();
  }
  return true;
}

// Strength graph: j != i with |a_ij| >= theta sqrt(a_ii a_jj) (1/dinv = a_ii).
// kStrengthLanes lanes per row take its entries in groups (lane t the t-th of
// a group); the fill pass keeps the row order through the group's ballot.
constexpr int kStrengthLanes = 4;

template <bool kFill>
__global__ void strength_kernel(int64_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                const double* __restrict__ v, const double* __restrict__ dinv, double theta,
                                int32_t* __restrict__ cnt, const int32_t* __restrict__ gp, int32_t* __restrict__ gj) {
  constexpr int G = kStrengthLanes;
  const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
  const int t = threadIdx.x & (G - 1);
  const int gshift = (threadIdx.x & 63) & ~(G - 1);  // first lane of this row's group
  if (i >= n) return;  // the whole group
  const double di = dinv[i];
  const int s0 = rp[i], e0 = rp[i + 1];
  int c = 0;
  int o = kFill ? gp[i] : 0;
  for (int q0 = s0; q0 < e0; q0 += G) {
    const int q = q0 + t;
    bool pass = false;
    int j = 0;
    if (q < e0) {
      j = ci[q];
      // |a_ij|^2 >= theta^2 a_ii a_jj  <=>  |a_ij|^2 dinv_i dinv_j >= theta^2
      const double a = v[q];
      pass = j != (int)i && a * a * di * dinv[j] >= theta * theta;
    }
    const unsigned grp = (unsigned)(__ballot(pass) >> gshift) & ((1u << G) - 1u);
    if (kFill && pass) gj[o + __popc(grp & ((1u << t) - 1u))] = j;
    o += __popc(grp);
    c += __popc(grp);
  }
  if (!kFill && t == 0) cnt[i] = c;
}

void strength_graph(const Csr& A, const double* dinv, double theta, Csr& G, hipStream_t s) {
  const int64_t n = A.n;
  DevBuf<int32_t> cnt(n > 0 ? n : 1);
  const unsigned grid = grid_for(n * kStrengthLanes, 256);
  if (n > 0) {
    strength_kernel<false><<<grid, 256, 0, s>>>(n, A.rp.get(), A.ci.get(), A.v64.get(), dinv, theta, cnt.get(),
                                                nullptr, nullptr);
    HIP_LAUNCH_CHECK();
  }
  G.n = G.m = n;
  G.rp.resize(n + 1);
  G.nnz = dev_exclusive_scan_i32(cnt.get(), G.rp.get(), n, s);
  G.ci.resize(G.nnz > 0 ? G.nnz : 1);
  if (n > 0) {
    strength_kernel<true><<<grid, 256, 0, s>>>(n, A.rp.get(), A.ci.get(), A.v64.get(), dinv, theta, nullptr,
                                               G.rp.get(), G.ci.get());
    HIP_LAUNCH_CHECK();
  }
}

void to_f32(Csr& A, hipStream_t s) {
  A.v32.resize(A.nnz > 0 ? A.nnz : 1);
  dev_f64_to_f32(A.v64.get(), A.v32.get(), A.nnz, s);
}

}  // namespace

void aggregate_graph(int64_t n, const int32_t* adj_ptr, const int32_t* adj, int depth, int min_size, int rule,
                     unsigned long long seed, int32_t* agg_out, int64_t* n_agg, hipStream_t s) {
  if (n == 0) {
    *n_agg = 0;
    return;
  }
  DevBuf<int> mis(n);
  randomized_mis(n, adj_ptr, adj, depth < 1 ? 1 : depth, seed, mis.get(), s);
  // roots numbered by rank (aggregateGraph: inclusive scan + ifLabelOne)
  DevBuf<int32_t> scan(n + 1);
  DevBuf<int> part_in(n), part_out(n), aggregated(n);
  // mis is 0/1 after completion, so its exclusive scan numbers the roots
  dev_exclusive_scan_i32(mis.get(), scan.get(), n, s);
  roots_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, mis.get(), scan.get(), part_in.get(), aggregated.get());
  HIP_LAUNCH_CHECK();
  DevBuf<unsigned> cnt(1);
  DevBuf<int> sizes;
  unsigned prev_unalloc = 0xffffffffu;
  for (int loop = 0;; ++loop) {
    if (loop > 1000000) fail(FEM_ERR_NUMERIC, "aggregation did not complete");
    allocate_kernel<<<grid_for(n, kAllocNT), kAllocNT, 0, s>>>(n, rule, adj_ptr, adj, part_in.get(), part_out.get(),
                                                               aggregated.get());
    HIP_LAUNCH_CHECK();
    std::swap(part_in, part_out);
    HIP_CHECK(hipMemsetAsync(cnt.get(), 0, sizeof(unsigned), s));
    count_unallocated_kernel<<<reduce_grid(n), 256, 0, s>>>(n, aggregated.get(), cnt.get());
    HIP_LAUNCH_CHECK();
    unsigned unalloc = read_counter(cnt, s);
    if (unalloc == 0) {
      // removeRuntyParts
      DevBuf<int> mx(1);
      dev_fill_i32(mx.get(), -1, 1, s);
      max_kernel<<<reduce_grid(n), 256, 0, s>>>(n, part_in.get(), mx.get());
      HIP_LAUNCH_CHECK();
      int hmx = -1;
      HIP_CHECK(hipMemcpyAsync(&hmx, mx.get(), sizeof(int), hipMemcpyDeviceToHost, s));
      HIP_CHECK(hipStreamSynchronize(s));
      int64_t np = hmx + 1;
      sizes.resize(np);
      dev_fill_i32(sizes.get(), 0, np, s);
      part_sizes_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, part_in.get(), sizes.get());
      HIP_LAUNCH_CHECK();
      DevBuf<int32_t> st(np), excl(np + 1);
      runt_stencil_kernel<<<grid_for(np, 256), 256, 0, s>>>(np, sizes.get(), min_size, st.get());
      HIP_LAUNCH_CHECK();
      int64_t removed = dev_exclusive_scan_i32(st.get(), excl.get(), np, s);
      if (removed == 0) {
        *n_agg = np;
        break;
      }
      runt_apply_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, part_in.get(), st.get(), excl.get(), aggregated.get());
      HIP_LAUNCH_CHECK();
      prev_unalloc = 0xffffffffu;
      continue;
    }
    if (unalloc == prev_unalloc) {
      // Nodes with no aggregated node in reach (a component left without a
      // root after runt removal): each becomes its own aggregate and the
      // aggregation ends (the reference loops forever here).
      DevBuf<int> mx(1);
      dev_fill_i32(mx.get(), -1, 1, s);
      max_kernel<<<reduce_grid(n), 256, 0, s>>>(n, part_in.get(), mx.get());
      HIP_LAUNCH_CHECK();
      int hmx = -1;
      HIP_CHECK(hipMemcpyAsync(&hmx, mx.get(), sizeof(int), hipMemcpyDeviceToHost, s));
      HIP_CHECK(hipStreamSynchronize(s));
      DevBuf<int32_t> fl(n), ex(n + 1);
      stuck_flag_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, aggregated.get(), fl.get());
      HIP_LAUNCH_CHECK();
      int64_t nst = dev_exclusive_scan_i32(fl.get(), ex.get(), n, s);
      stuck_assign_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, hmx + 1, fl.get(), ex.get(), part_in.get());
      HIP_LAUNCH_CHECK();
      *n_agg = hmx + 1 + nst;
      break;
    }
    prev_unalloc = unalloc;
  }
  HIP_CHECK(hipMemcpyAsync(agg_out, part_in.get(), sizeof(int32_t) * n, hipMemcpyDeviceToDevice, s));
  HIP_CHECK(hipStreamSynchronize(s));
}

// ---- precision 2: level-0 smoother operator D^-1/2 A D^-1/2 with FP16 off-diagonals

// off-diagonal part of row i, scaled: a_ij s_i s_j with s = 1/sqrt(diag) (the
// product s_i s_j is formed first so that entries (i,j) and (j,i) are equal)
__global__ void offdiag_scaled_kernel(int64_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                      const double* __restrict__ v, const double* __restrict__ dinv,
                                      int32_t* __restrict__ orp, int32_t* __restrict__ oci,
                                      double* __restrict__ ov) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  orp[i] = rp[i] - (int32_t)i;  // every row holds its diagonal once
  if (i == n) return;
  const double si = sqrt(dinv[i]);
  int o = rp[i] - (int32_t)i;
  for (int q = rp[i]; q < rp[i + 1]; ++q) {
    const int j = ci[q];
    if (j == (int)i) continue;
    oci[o] = j;
    ov[o] = v[q] * (si * sqrt(dinv[j]));
    ++o;
  }
}

// diagonal 1 + eps_i that keeps the row sums of A:
// eps_i = - sum_j (fp16(a^_ij) - a^_ij) sqrt(a_jj / a_ii); and s_i = 1/sqrt(a_ii)
__global__ void half_diag_kernel(int64_t n, const int32_t* __restrict__ orp, const int32_t* __restrict__ oci,
                                 const double* __restrict__ ov, const double* __restrict__ dinv,
                                 float* __restrict__ diag, float* __restrict__ sv) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double si = sqrt(dinv[i]);
  double eps = 0.0;
  for (int q = orp[i]; q < orp[i + 1]; ++q) {
    const double a = ov[q];
    eps -= (half_value(a) - a) * (si / sqrt(dinv[oci[q]]));
  }
  diag[i] = (float)(1.0 + eps);
  sv[i] = (float)si;
}

// rows of P times sqrt(a_ii)
__global__ void scale_rows_sqrt_kernel(int64_t n, const int32_t* __restrict__ rp, const double* __restrict__ dinv,
                                       double* __restrict__ v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double q = 1.0 / sqrt(dinv[i]);
  for (int p = rp[i]; p < rp[i + 1]; ++p) v[p] *= q;
}

void copy_csr(const Csr& A, Csr& B, hipStream_t s) {
  B.n = A.n;
  B.m = A.m;
  B.nnz = A.nnz;
  B.rp.resize(A.n + 1);
  B.ci.resize(A.nnz > 0 ? A.nnz : 1);
  B.v64.resize(A.nnz > 0 ? A.nnz : 1);
  HIP_CHECK(hipMemcpyAsync(B.rp.get(), A.rp.get(), sizeof(int32_t) * (A.n + 1), hipMemcpyDeviceToDevice, s));
  HIP_CHECK(hipMemcpyAsync(B.ci.get(), A.ci.get(), sizeof(int32_t) * A.nnz, hipMemcpyDeviceToDevice, s));
  HIP_CHECK(hipMemcpyAsync(B.v64.get(), A.v64.get(), sizeof(double) * A.nnz, hipMemcpyDeviceToDevice, s));
}

namespace {

// Lanes per row of a sliced-ELL copy, from the average row length: one lane
// up to 24 entries (fine operators, prolongators), else the fewest lanes that
// leave at most 20 entries to each (161^3 level 1 without LDS staging: its
// ~46-entry rows take 4 lanes, 18.4 us per Chebyshev step against 34.6 us for
// the 16-lane CSR kernel; ~140-entry level-0 restriction rows take 8, 37.4 us against 62.4 us;
// 2x lanes, or rows sorted by length, measured slower).  0 = no sliced copy:
// operators under 32768 rows with long rows have too few slices to fill the
// chip and keep the CSR kernel (4,904 rows: 7.2 us sliced, 4.8 us CSR).
constexpr int64_t kLdsCoarseRows = 65536;

// slices per LDS staging block: 8 (512 rows) where that still gives at least
// 2048 blocks, else 2 so that a few thousand rows still fill the 256 CUs
int lds_block_slices(const Csr& M) { return M.nslices >= 8 * 2048 ? 8 : 2; }

int sell_lanes(const Csr& M) {
  if (M.n == 0) return 0;
  const double avg = (double)M.nnz / (double)M.n;
  if (avg <= 24.0) return 1;
  if (M.n < 32768) return 0;
  int g = 2;
  while (g < 32 && avg > 20.0 * g) g *= 2;
  return g;
}

void sell_or_csr(Csr& M, int64_t sort_window, int lanes, bool with_f32, hipStream_t s, bool split_diag = false) {
  if (lanes > 0) {
    build_sell(M, with_f32, s, sort_window, false, true, lanes, split_diag && lanes == 1);
    return;
  }
  M.nslices = 0;
  M.sell_g = 1;
  M.srow.release();
  M.sp.release();
  M.sc.release();
  M.sc16.release();
  M.swin.release();
  M.s64.release();
  M.s32.release();
  M.sd64.release();
  M.sd32.release();
  M.nnz_sdiag = 0;
  M.lds_bs = 0;
  M.lds_maxc = 0;
  M.lc.release();
  M.lcs.release();
  M.lcols.release();
}

}  // namespace

void build_half_level0(Handle& h) {
  hipStream_t s = h.stream;
  Hierarchy& H = h.H;
  Level& L0 = *H.levels[0];
  const Csr& A = h.A;
  const int64_t n = A.n;
  Csr& O = H.ah;
  O.n = O.m = n;
  O.nnz = A.nnz - n;
  O.rp.resize(n + 1);
  O.ci.resize(O.nnz > 0 ? O.nnz : 1);
  O.v64.resize(O.nnz > 0 ? O.nnz : 1);
  offdiag_scaled_kernel<<<grid_for(n + 1, 256), 256, 0, s>>>(n, A.rp.get(), A.ci.get(), A.v64.get(),
                                                              L0.dinv64.get(), O.rp.get(), O.ci.get(), O.v64.get());
  HIP_LAUNCH_CHECK();
  H.ah_diag.resize(n);
  H.ah_s.resize(n);
  half_diag_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, O.rp.get(), O.ci.get(), O.v64.get(), L0.dinv64.get(),
                                                     H.ah_diag.get(), H.ah_s.get());
  HIP_LAUNCH_CHECK();
  build_sell(O, false, s, 0, /*with_f16=*/true, /*with_f64=*/false);
  // y staged in LDS per block of slices, as for the FP32 fine operator
  build_lds(O, lds_block_slices(O), s);
  O.v64.release();  // only the FP16 sliced copy is used
  // D^1/2 P (sliced ELL, FP32, rows sorted by length) and its transpose (CSR, FP32)
  copy_csr(L0.P, H.ph, s);
  scale_rows_sqrt_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, H.ph.rp.get(), L0.dinv64.get(), H.ph.v64.get());
  HIP_LAUNCH_CHECK();
  csr_transpose(H.ph, H.rh, s);
  to_f32(H.ph, s);
  to_f32(H.rh, s);
  build_sell(H.ph, true, s, 1024);
  sell_or_csr(H.rh, 0, sell_lanes(H.rh), true, s);
}

void build_hierarchy(Handle& h) {
  auto t0 = std::chrono::steady_clock::now();
  ++h.gen;  // invalidates a captured PCG iteration
  hipStream_t s = h.stream;
  Hierarchy& H = h.H;
  H.levels.clear();
  H.cinv64.release();
  H.precision = h.prm.precision;
  H.built = false;
  const int fine_depth = h.prm.randMisParameters % 100;
  const int min_size = (h.prm.randMisParameters / 10000) % 10;
  const double omega = h.prm.proOmega;
  // aggregatorType 3 (AggMIS "NewMIS" fine aggregation,
  // src/core/cuda/ComputePermutationMethods.cu:352-431): its own reading of
  // randMisParameters -- sizes [p / 1000 % 1000, p % 1000], MIS depth
  // p / 10^7 % 10 -- then MIS, aggregate-to-nearest and the merge-split
  // conditioner.  Depth 0 (the default's digit) is taken as 2: the
  // reference's depth-0 MIS marks no root and its aggregate-to-nearest loop
  // never ends.
  // aggregatorType 4 and 5 (NewMIS_CPU, LightMIS_CPU, ComputePermutationMethods.cu:566-775 and
  // :776-860): the same reading, a flood-fill MIS of depth p / 10^7 % 10 on the
  // host (0 is a plain MIS there), aggregation to the nearest root, and for
  // type 4 the conditioner (type 5 builds one and does not run it).
  const bool aggmis = h.prm.aggregatorType == 3;
  const bool flood = h.prm.aggregatorType == 4 || h.prm.aggregatorType == 5;
  const int agg_lo = (h.prm.randMisParameters / 1000) % 1000, agg_hi = h.prm.randMisParameters % 1000;
  const int flood_depth = (h.prm.randMisParameters / 10000000) % 10;
  const int agg_depth = flood_depth ? flood_depth : 2;

  // level-0 graph: pattern of A without the diagonal
  Csr G;
  if (h.prm.strengthThreshold <= 0.0) offdiag_graph(h.A, G, s);
  // the fine operator's sliced copy first (the power iteration below runs on
  // it), with its diagonal outside the slices: the diagonal term then needs
  // no column code and no gather (x_i is read by the row anyway; measured
  // 3-5 % per level-0 pass at 161^3)
  sell_or_csr(h.A, 0, sell_lanes(h.A), H.precision >= 1, s, true);
  // blocks of 8 slices stage their distinct columns of x in LDS (measured at
  // 161^3: 8 slices per block 75 us per FP32 SpMV against 77 with 16 and 94
  // with gathers from memory)
  build_lds(h.A, lds_block_slices(h.A), s);

  int64_t total_rows = 0, total_nnz = 0;
  std::unique_ptr<Level> cur(new Level());
  for (int l = 0;; ++l) {
    Level& L = *cur;
    const Csr& A = (l == 0) ? h.A : L.A;
    const int64_t n = A.n;
    total_rows += n;
    total_nnz += A.nnz;
    L.dinv64.resize(n);
    {
      DevBuf<int> bad(1);
      dev_fill_i32(bad.get(), 0, 1, s);
      diag_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, A.rp.get(), A.ci.get(), A.v64.get(), L.dinv64.get(), bad.get());
      HIP_LAUNCH_CHECK();
      int hb = 0;
      HIP_CHECK(hipMemcpyAsync(&hb, bad.get(), sizeof(int), hipMemcpyDeviceToHost, s));
      HIP_CHECK(hipStreamSynchronize(s));
      if (hb) fail(FEM_ERR_NUMERIC, "level " + std::to_string(l) + " has " + std::to_string(hb) +
                                        " rows without a positive diagonal");
    }
    const bool last = (n < h.prm.topSize) || (l + 1 >= h.prm.maxLevels) || n <= 1;
    int64_t nagg = 0;
    const double theta = h.prm.strengthThreshold;
    if (!last && theta > 0.0) strength_graph(A, L.dinv64.get(), theta, G, s);
    if (!last) {
      L.agg.resize(n);
      unsigned long long seed = h.prm.seed + 0x9E37ull * (unsigned long long)l;
      auto aggregate = [&]() {
        if (flood) {
          std::vector<int32_t> grp(n + 1), gci(G.nnz > 0 ? G.nnz : 1), hagg;
          G.rp.download(grp.data(), n + 1, s);
          if (G.nnz > 0) G.ci.download(gci.data(), G.nnz, s);
          HIP_CHECK(hipStreamSynchronize(s));
          const std::vector<int32_t> roots =
              flood_fill_mis(n, grp.data(), gci.data(), flood_depth, flood_fill_starter(n, seed));
          nagg = aggregate_to_nearest(n, grp.data(), gci.data(), roots, hagg);
          if (h.prm.aggregatorType == 4) nagg = condition_aggregates(n, grp.data(), gci.data(), hagg, nagg, agg_lo, agg_hi);
          L.agg.upload(hagg.data(), n, s);
          return;
        }
        if (!aggmis) {
          aggregate_graph(n, G.rp.get(), G.ci.get(), fine_depth, min_size, h.prm.aggregationRule, seed, L.agg.get(),
                          &nagg, s);
          return;
        }
        // MIS + aggregate-to-nearest on the device (AggregateToNearest,
        // src/core/aggmis/cuda/AggMIS_Aggregation_GPU.cu:114-158: its
        // candidate table makes it the first-aggregated-neighbour rule, no
        // runt removal), then the conditioner on the host
        aggregate_graph(n, G.rp.get(), G.ci.get(), agg_depth, 0, 0, seed, L.agg.get(), &nagg, s);
        std::vector<int32_t> grp(n + 1), gci(G.nnz), hagg(n);
        G.rp.download(grp.data(), n + 1, s);
        G.ci.download(gci.data(), G.nnz, s);
        L.agg.download(hagg.data(), n, s);
        HIP_CHECK(hipStreamSynchronize(s));
        nagg = condition_aggregates(n, grp.data(), gci.data(), hagg, nagg, agg_lo, agg_hi);
        L.agg.upload(hagg.data(), n, s);
      };
      aggregate();
      if (theta > 0.0 && 2 * nagg > n) {
        // The strength graph fell apart into small components (each left as
        // singletons after runt removal): coarsening would stall for many
        // levels.  Aggregate this level on the full pattern instead.
        offdiag_graph(A, G, s);
        aggregate();
      }
    }
    if (last || nagg >= n || nagg == 0) {
      // coarsest level: dense FP64 inverse
      if (n > kMaxCoarse)
        fail(FEM_ERR_CAPACITY, "coarsest level has " + std::to_string(n) + " rows (limit " +
                                   std::to_string(kMaxCoarse) + "); raise maxLevels or lower topSize");
      L.agg.release();
      std::vector<int32_t> rp(n + 1), ci(A.nnz);
      std::vector<double> v(A.nnz);
      A.rp.download(rp.data(), n + 1, s);
      A.ci.download(ci.data(), A.nnz, s);
      A.v64.download(v.data(), A.nnz, s);
      HIP_CHECK(hipStreamSynchronize(s));
      std::vector<double> dense(n * n, 0.0);
      for (int64_t i = 0; i < n; ++i)
        for (int q = rp[i]; q < rp[i + 1]; ++q) dense[i * n + ci[q]] += v[q];
      if (!dense_inverse(dense, n)) fail(FEM_ERR_NUMERIC, "coarsest-level matrix is singular");
      // applied in FP64 in every precision mode (coarse_kernel)
      H.cinv64.upload(dense.data(), n * n, s);
      H.coarse = l;
      H.nc = n;
      H.levels.push_back(std::move(cur));
      break;
    }
    L.nc = nagg;
    L.rho = estimate_rho(A, L.dinv64.get(), s);
    // prolongator pattern: columns agg(k) for k in row i of A
    if (!agg_pattern(A, L.agg.get(), nagg, L.P, s))
      spgemm_symbolic(A, nullptr, nullptr, 0, L.agg.get(), nagg, false, L.P, s);
    L.P.v64.resize(L.P.nnz);
    prolongator_kernel<<<grid_for(n, kProNT), kProNT, 0, s>>>(n, A.rp.get(), A.ci.get(), A.v64.get(),
                                                              L.dinv64.get(), L.agg.get(), omega, L.P.rp.get(),
                                                              L.P.ci.get(), L.P.v64.get());
    HIP_LAUNCH_CHECK();
    csr_transpose(L.P, L.R, s);
    // A_c = R (A P)
    std::unique_ptr<Level> next(new Level());
    {
      Csr AP;
      spgemm_symbolic(A, L.P.rp.get(), L.P.ci.get(), L.P.n, nullptr, nagg, false, AP, s);
      spgemm_numeric(A, L.P, AP, s);
      spgemm_symbolic(L.R, AP.rp.get(), AP.ci.get(), AP.n, nullptr, nagg, false, next->A, s);
      spgemm_numeric(L.R, AP, next->A, s);
    }
    // next graph: induced aggregate graph (aggregates sharing an edge); with a
    // strength threshold the next level builds its own strength graph instead
    if (theta <= 0.0) {
      Csr M;
      aggregate_members(L.agg.get(), n, nagg, M, s);
      Csr G2;
      spgemm_symbolic(M, G.rp.get(), G.ci.get(), G.n, L.agg.get(), nagg, true, G2, s);
      G = std::move(G2);
    }
    H.levels.push_back(std::move(cur));
    cur = std::move(next);
  }
  H.grid_complexity = (double)total_rows / (double)h.A.n;
  H.operator_complexity = (double)total_nnz / (double)h.A.nnz;
  if (H.precision >= 1) {
    for (size_t l = 0; l < H.levels.size(); ++l) {
      Level& L = *H.levels[l];
      Csr& A = (l == 0) ? h.A : L.A;
      to_f32(A, s);
      L.dinv32.resize(L.dinv64.size());
      dev_f64_to_f32(L.dinv64.get(), L.dinv32.get(), L.dinv64.size(), s);
      if ((int)l != H.coarse) {
        to_f32(L.P, s);
        to_f32(L.R, s);
      }
    }
  }
  // Sliced-ELL copies for the solve phase, built for every operator, restriction and
  // prolongator the one-column solve applies, except small coarse operators with long rows
  // (see sell_lanes).
  for (size_t l = 0; l < H.levels.size(); ++l) {
    Level& L = *H.levels[l];
    Csr& A = (l == 0) ? h.A : L.A;
    // operators stay in row order (the fused PCG and level-0 smoother kernels
    // read the fine one with one lane per row, in row order; built above).
    // Coarse operators of at least kLdsCoarseRows rows take one lane per row
    // with LDS staging in blocks of 2 slices (161^3 level 1, 181,696 rows of
    // ~46 entries: Chebyshev step 14.6 us, against 18.0 us with 8-slice
    // blocks (355 blocks for 256 CUs), 15.5 us with 1-slice blocks and
    // 18.4 us for the 4-lane copy without staging)
    if (l > 0 && A.n >= kLdsCoarseRows) {
      sell_or_csr(A, 0, 1, H.precision >= 1, s, true);
      build_lds(A, lds_block_slices(A), s);
    } else if (l > 0) {
      sell_or_csr(A, 0, sell_lanes(A), H.precision >= 1, s);
    }
    if ((int)l != H.coarse) {
      // prolongator rows (1 to ~10 entries) are sorted by length within windows
      // of 1024 rows: 1.8x padding in row order at 81^3, ~1.0x sorted
      sell_or_csr(L.P, 1024, sell_lanes(L.P), H.precision >= 1, s);
      // prolongators of many rows: the coarse vector staged in LDS (161^3 level
      // 0: 52.5 -> 41.0 us per prolongation)
      if (L.P.n >= kLdsCoarseRows) build_lds(L.P, lds_block_slices(L.P), s);
      sell_or_csr(L.R, 0, sell_lanes(L.R), H.precision >= 1, s);
    }
  }
  H.ah.nslices = 0;
  H.ph.nslices = 0;
  H.rh.nslices = 0;
  if (H.precision == 2 && H.coarse > 0) build_half_level0(h);
  HIP_CHECK(hipStreamSynchronize(s));
  H.built = true;
  H.setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace fem
