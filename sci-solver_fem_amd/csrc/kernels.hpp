// Device kernels of the solve phase: CSR SpMV / SpMM with fused smoother
// epilogues, reductions with a last-block finish, and the dense coarse solve.
//
// Layout: vectors with k right-hand sides are row-major (n x k), so the k
// values of one vertex are contiguous and a gather x[col*k + c] by k lanes is
// one coalesced access.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "fem_internal.hpp"

namespace fem {

constexpr int kNT = 256;  // threads per block of the solve kernels

enum SpmvMode : int {
  kAssign = 0,   // y = A x
  kResid = 1,    // y = b - A x
  kAdd = 2,      // y += A x
  kJacobi = 3,   // xn = x + c1 * dinv * (b - A x)
  kCheb = 4,     // r = dinv (b - A x); d = c1 d + c2 r; xn = x + d
  kAssignScale = 5,  // y = A x; xn = c1 * dinv * y; d = xn when d != null (a
                     // restriction that also takes the next level's first
                     // smoothing step from zero)
};

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  Gathering kernels remap the hardware block index so that XCD j runs a
// contiguous range of logical blocks: neighbouring rows (Morton order) then
// share one L2 and the gathered vector is fetched from HBM about once
// instead of once per XCD.  A bijection of [0, nblocks) for any nblocks.
constexpr unsigned kNumXcd = 8;

__device__ __forceinline__ unsigned xcd_block(unsigned bid, unsigned nblocks) {
  const unsigned x = bid % kNumXcd, i = bid / kNumXcd;
  const unsigned q = nblocks / kNumXcd, r = nblocks % kNumXcd;
  return x * q + min(x, r) + i;
}

// Rows of a block in a loop over n rows, `per` rows per block and pass, laid
// out so that each XCD sweeps one contiguous range of rows with a single
// front: the blocks of one XCD take consecutive groups of `per` rows of its
// range, then all advance together by (its blocks) x per.  Contiguous chunks
// per block instead give each XCD hundreds of separate fronts whose gathered
// neighbourhoods do not fit its L2 (measured for 16 FP64 columns at 161^3:
// 3.2x the algorithmic HBM reads).  The XCD ranges are those of the
// contiguous logical blocks xcd_block assigns to it, so the rows are covered
// once for any grid size.
struct XcdRows {
  int64_t hi;     // end of this XCD's rows
  int64_t first;  // first row group of this block
  int64_t step;   // rows between its groups
};

__device__ __forceinline__ XcdRows xcd_rows(int64_t n, int64_t per) {
  const unsigned nb = gridDim.x, bid = blockIdx.x;
  const unsigned x = bid % kNumXcd, i = bid / kNumXcd;
  const unsigned q = nb / kNumXcd, r = nb % kNumXcd;
  const unsigned nbx = q + (x < r ? 1u : 0u);  // blocks on this XCD
  const unsigned lb0 = x * q + min(x, r);      // their first logical block
  const int64_t chunk = ((n + nb - 1) / nb + per - 1) / per * per;
  const int64_t lo = (int64_t)lb0 * chunk;
  XcdRows R;
  R.hi = min(n, (int64_t)(lb0 + nbx) * chunk);
  R.first = lo + (int64_t)i * per;
  R.step = (int64_t)nbx * per;
  return R;
}

// Matrix entries are read once per SpMV: nontemporal loads keep them from
// evicting the gathered vector from L2 (measured: -15 % time at 4.2M rows).
template <typename T>
__device__ __forceinline__ T ntload(const T* p) {
  return __builtin_nontemporal_load(p);
}

template <typename T>
__device__ __forceinline__ T shfl_xor_t(T v, int mask, int width) {
  return __shfl_xor(v, mask, width);
}

// One epilogue for all modes; `acc` = (A x)_row for column c.
template <int MODE, typename T, typename V>
__device__ __forceinline__ void spmv_epilogue(int64_t idx, int64_t row, T acc, T* __restrict__ y,
                                              const T* __restrict__ x, const T* __restrict__ b,
                                              const V* __restrict__ dinv, T* __restrict__ d,
                                              T* __restrict__ xn, T c1, T c2) {
  if (MODE == kAssign) {
    y[idx] = acc;
  } else if (MODE == kAssignScale) {
    y[idx] = acc;
    const T v = c1 * (T)dinv[row] * acc;
    xn[idx] = v;
    if (d) d[idx] = v;
  } else if (MODE == kResid) {
    y[idx] = b[idx] - acc;
  } else if (MODE == kAdd) {
    y[idx] += acc;
  } else if (MODE == kJacobi) {
    xn[idx] = x[idx] + c1 * (T)dinv[row] * (b[idx] - acc);
  } else {  // kCheb: b, D^-1 and d are read once; nontemporal for FP64 (measured faster there, not for FP32)
    constexpr bool nt = sizeof(T) == 8;
    T r = (T)(nt ? ntload(dinv + row) : dinv[row]) * ((nt ? ntload(b + idx) : b[idx]) - acc);
    T dn = (c1 != T(0) ? c1 * (nt ? ntload(d + idx) : d[idx]) : T(0)) + c2 * r;
    d[idx] = dn;
    xn[idx] = x[idx] + dn;
  }
}

// Single right-hand side: G lanes per row, strided over the row's entries,
// reduced by a butterfly (fixed order).
template <typename V, typename T, int G, int MODE>
__global__ void __launch_bounds__(kNT) csr_vec_kernel(int64_t n, const int32_t* __restrict__ rp,
                                                    const int32_t* __restrict__ ci, const V* __restrict__ val,
                                                    const T* __restrict__ x, T* __restrict__ y,
                                                    const T* __restrict__ b, const V* __restrict__ dinv,
                                                    T* __restrict__ d, T* __restrict__ xn, T c1, T c2) {
  const int64_t row = ((int64_t)xcd_block(blockIdx.x, gridDim.x) * kNT + threadIdx.x) / G;
  const int lane = threadIdx.x & (G - 1);
  if (row >= n) return;
  const int s = rp[row], e = rp[row + 1];
  T acc = T(0);
  for (int k = s + lane; k < e; k += G) acc += (T)ntload(val + k) * x[ntload(ci + k)];
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) acc += shfl_xor_t(acc, off, G);
  if (lane == 0) spmv_epilogue<MODE, T, V>(row, row, acc, y, x, b, dinv, d, xn, c1, c2);
}

// k right-hand sides: KP lanes per row (KP = power of two >= k), lane c owns
// column c and walks the row sequentially.
template <typename V, typename T, int KP, int MODE>
__global__ void __launch_bounds__(kNT) csr_multi_kernel(int64_t n, int k, const int32_t* __restrict__ rp,
                                                      const int32_t* __restrict__ ci, const V* __restrict__ val,
                                                      const T* __restrict__ x, T* __restrict__ y,
                                                      const T* __restrict__ b, const V* __restrict__ dinv,
                                                      T* __restrict__ d, T* __restrict__ xn, T c1, T c2) {
  const int64_t row = ((int64_t)xcd_block(blockIdx.x, gridDim.x) * kNT + threadIdx.x) / KP;
  const int c = threadIdx.x & (KP - 1);
  if (row >= n || c >= k) return;
  const int s = rp[row], e = rp[row + 1];
  T acc = T(0);
  // plain loads: the KP lanes of a row walk its entries one by one, so each
  // cache line of val / ci is used over several iterations (a nontemporal
  // hint would evict it after the first).  Four entries per step keep four
  // independent gathers in flight; the sum stays in entry order.
  int q = s;
  for (; q + 4 <= e; q += 4) {
    const int c0 = ci[q], c1i = ci[q + 1], c2i = ci[q + 2], c3 = ci[q + 3];
    const T v0 = (T)val[q], v1 = (T)val[q + 1], v2 = (T)val[q + 2], v3 = (T)val[q + 3];
    const T g0 = x[(int64_t)c0 * k + c], g1 = x[(int64_t)c1i * k + c], g2 = x[(int64_t)c2i * k + c],
            g3 = x[(int64_t)c3 * k + c];
    acc += v0 * g0;
    acc += v1 * g1;
    acc += v2 * g2;
    acc += v3 * g3;
  }
  for (; q < e; ++q) acc += (T)val[q] * x[(int64_t)ci[q] * k + c];
  spmv_epilogue<MODE, T, V>(row * k + c, row, acc, y, x, b, dinv, d, xn, c1, c2);
}

// Sliced ELL, one RHS: a wavefront owns a slice of 64 rows, lane = row; the
// slice is column-major so every step k is one coalesced 64-wide load of
// values and of columns, and the k steps are independent loads the compiler
// keeps in flight together.  Sequential per-row sum (fixed order).
struct SellView {
  const int32_t* sp;    // nslices + 1 element offsets
  const int32_t* sc;    // int32 columns
  const uint16_t* sc16; // 16-bit column codes of compact slices
  const int32_t* swin;  // kSellWindows window bases per slice (-1: not compact)
  const int32_t* srow;  // row at each position (null: position = row)
  const double* sd64;   // diagonal outside the slices (null: in the slices)
  const float* sd32;
};

inline SellView sell_view(const Csr& A) {
  return SellView{A.sp.get(), A.sc.get(), A.sc16.get(), A.swin.get(), A.srow.size() ? A.srow.get() : nullptr,
                  A.sd64.size() ? A.sd64.get() : nullptr, A.sd32.size() ? A.sd32.get() : nullptr};
}

// the diagonal held outside the slices, in the working precision (null when
// the slices hold it)
template <typename T>
__device__ __forceinline__ const T* sell_diag(const SellView& S);
template <>
__device__ __forceinline__ const double* sell_diag<double>(const SellView& S) { return S.sd64; }
template <>
__device__ __forceinline__ const float* sell_diag<float>(const SellView& S) { return S.sd32; }

// Window bases of a compact slice, read once per wavefront into scalar
// registers (named members, not an array: the select tree below must stay
// register selects).
struct SellWin {
  int b0, b1, b2, b3, b4, b5, b6, b7;
};

__device__ __forceinline__ int sell_col(unsigned c, SellWin W) {
  const bool s0 = c & (1u << kSellWinBits), s1 = c & (2u << kSellWinBits), s2 = c & (4u << kSellWinBits);
  const int lo = s1 ? (s0 ? W.b3 : W.b2) : (s0 ? W.b1 : W.b0);
  const int hi = s1 ? (s0 ? W.b7 : W.b6) : (s0 ? W.b5 : W.b4);
  return (s2 ? hi : lo) + (int)(c & ((1u << kSellWinBits) - 1u));
}

// value of a stored matrix entry: FP64 / FP32 as stored, FP16 kept as its
// 16-bit pattern (uint16_t)
template <typename T, typename V>
__device__ __forceinline__ T sell_val(const V* p) {
  return (T)ntload(p);
}
template <>
__device__ __forceinline__ float sell_val<float, uint16_t>(const uint16_t* p) {
  return (float)__builtin_bit_cast(_Float16, ntload(p));
}

// Row sum of slice `slice` (wave-uniform) for this lane's row.  Compact
// slices read 2-byte column codes: bits 13-15 select one of the slice's 8
// window bases, bits 0-12 are the offset.
template <typename T, typename V = T>
__device__ __forceinline__ T sell_row_sum(const SellView& S, int slice, int lane, const V* __restrict__ sv,
                                          const T* __restrict__ x) {
  const int off = __builtin_amdgcn_readfirstlane(S.sp[slice]);
  const int w = __builtin_amdgcn_readfirstlane((S.sp[slice + 1] - off) >> 6);
  const V* vp = sv + off + lane;
  T acc = T(0);
  static_assert(kSellWindows == 8, "sell_col selects among 8 windows");
  const int32_t* wp = S.swin + slice * kSellWindows;
  const SellWin W{__builtin_amdgcn_readfirstlane(wp[0]), __builtin_amdgcn_readfirstlane(wp[1]),
                  __builtin_amdgcn_readfirstlane(wp[2]), __builtin_amdgcn_readfirstlane(wp[3]),
                  __builtin_amdgcn_readfirstlane(wp[4]), __builtin_amdgcn_readfirstlane(wp[5]),
                  __builtin_amdgcn_readfirstlane(wp[6]), __builtin_amdgcn_readfirstlane(wp[7])};
  if (W.b0 >= 0) {
    const uint16_t* cp = S.sc16 + off + lane;
    int k = 0;
    for (; k + 4 <= w; k += 4) {
      const unsigned c0 = ntload(cp + 64 * k), c1 = ntload(cp + 64 * (k + 1)), c2 = ntload(cp + 64 * (k + 2)),
                     c3 = ntload(cp + 64 * (k + 3));
      const T v0 = sell_val<T>(vp + 64 * k), v1 = sell_val<T>(vp + 64 * (k + 1)),
              v2 = sell_val<T>(vp + 64 * (k + 2)), v3 = sell_val<T>(vp + 64 * (k + 3));
      acc += v0 * x[sell_col(c0, W)];
      acc += v1 * x[sell_col(c1, W)];
      acc += v2 * x[sell_col(c2, W)];
      acc += v3 * x[sell_col(c3, W)];
    }
    for (; k < w; ++k) acc += sell_val<T>(vp + 64 * k) * x[sell_col(ntload(cp + 64 * k), W)];
  } else {
    const int32_t* cp = S.sc + off + lane;
    int k = 0;
    for (; k + 4 <= w; k += 4) {
      const int c0 = ntload(cp + 64 * k), c1 = ntload(cp + 64 * (k + 1)), c2 = ntload(cp + 64 * (k + 2)),
                c3 = ntload(cp + 64 * (k + 3));
      const T v0 = sell_val<T>(vp + 64 * k), v1 = sell_val<T>(vp + 64 * (k + 1)),
              v2 = sell_val<T>(vp + 64 * (k + 2)), v3 = sell_val<T>(vp + 64 * (k + 3));
      acc += v0 * x[c0];
      acc += v1 * x[c1];
      acc += v2 * x[c2];
      acc += v3 * x[c3];
    }
    for (; k < w; ++k) acc += sell_val<T>(vp + 64 * k) * x[ntload(cp + 64 * k)];
  }
  return acc;
}

// ---- LDS-staged sliced ELL (one lane per row, row order) ------------------------
// A block of L.bs consecutive slices first copies x at the distinct columns
// its rows use into LDS (one global gather per distinct column instead of
// one per stored entry: 0.18 per entry at 161^3 with 8 slices), then reads
// the entries' x from LDS through 16-bit indices.
struct LdsView {
  int bs;                // slices per block
  const uint16_t* lc;    // LDS index per stored entry
  const int32_t* lcs;    // block offsets into lcols
  const int32_t* lcols;  // distinct columns per block, ascending
};

inline LdsView lds_view(const Csr& A) { return LdsView{A.lds_bs, A.lc.get(), A.lcs.get(), A.lcols.get()}; }

// x at the block's distinct columns into xs (then a block barrier)
template <typename T>
__device__ __forceinline__ void lds_stage(const LdsView& L, unsigned blk, const T* __restrict__ x, T* xs) {
  const int c0 = L.lcs[blk], nc = L.lcs[blk + 1] - c0;
  for (int t = threadIdx.x; t < nc; t += blockDim.x) xs[t] = x[L.lcols[c0 + t]];
  __syncthreads();
}

// row sum of slice `slice` with x from LDS (same entry order as sell_row_sum)
template <typename T, typename V = T>
__device__ __forceinline__ T lds_row_sum(const SellView& S, const LdsView& L, int slice, int lane,
                                         const V* __restrict__ sv, const T* xs) {
  const int off = __builtin_amdgcn_readfirstlane(S.sp[slice]);
  const int w = __builtin_amdgcn_readfirstlane((S.sp[slice + 1] - off) >> 6);
  const V* vp = sv + off + lane;
  const uint16_t* cp = L.lc + off + lane;
  T acc = T(0);
  int k = 0;
  for (; k + 4 <= w; k += 4) {
    const unsigned i0 = ntload(cp + 64 * k), i1 = ntload(cp + 64 * (k + 1)), i2 = ntload(cp + 64 * (k + 2)),
                   i3 = ntload(cp + 64 * (k + 3));
    const T v0 = sell_val<T>(vp + 64 * k), v1 = sell_val<T>(vp + 64 * (k + 1)), v2 = sell_val<T>(vp + 64 * (k + 2)),
            v3 = sell_val<T>(vp + 64 * (k + 3));
    acc += v0 * xs[i0];
    acc += v1 * xs[i1];
    acc += v2 * xs[i2];
    acc += v3 * xs[i3];
  }
  for (; k < w; ++k) acc += sell_val<T>(vp + 64 * k) * xs[ntload(cp + 64 * k)];
  return acc;
}

// SpMV with the fused epilogues on the LDS-staged copy; grid = one block per
// L.bs slices, dynamic LDS = the largest distinct-column count x sizeof(T)
template <typename T, int MODE>
__global__ void __launch_bounds__(kNT) sell_lds_kernel(int64_t n, int64_t nslices, SellView S, LdsView L,
                                                     const T* __restrict__ sv, const T* __restrict__ x,
                                                     T* __restrict__ y, const T* __restrict__ b,
                                                     const T* __restrict__ dinv, T* __restrict__ d,
                                                     T* __restrict__ xn, T c1, T c2) {
  extern __shared__ unsigned char lds_raw[];
  T* xs = reinterpret_cast<T*>(lds_raw);
  const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  lds_stage<T>(L, blk, x, xs);
  const int lane = threadIdx.x & 63;
  const T* sd = sell_diag<T>(S);
  for (int q = threadIdx.x >> 6; q < L.bs; q += kNT / 64) {
    const int slice = __builtin_amdgcn_readfirstlane((int)(blk * L.bs + q));
    if (slice >= nslices) break;
    T acc = lds_row_sum<T>(S, L, slice, lane, sv, xs);
    const int64_t pos = (int64_t)slice * 64 + lane;
    if (pos < n) {
      const int64_t row = S.srow ? (int64_t)S.srow[pos] : pos;
      if (sd) acc += sd[row] * x[row];  // the diagonal term, last
      spmv_epilogue<MODE, T, T>(row, row, acc, y, x, b, dinv, d, xn, c1, c2);
    }
  }
}

// G lanes per row (slices of 64 / G rows): each lane sums its strided share
// of the row, then a butterfly over the G lanes adds the shares (fixed order).
template <typename T, int MODE, int G = 1>
__global__ void __launch_bounds__(kNT) sell_kernel(int64_t n, int64_t nslices, SellView S, const T* __restrict__ sv,
                                                 const T* __restrict__ x, T* __restrict__ y,
                                                 const T* __restrict__ b, const T* __restrict__ dinv,
                                                 T* __restrict__ d, T* __restrict__ xn, T c1, T c2) {
  // the slice index is the same on every lane of the wavefront
  const int slice =
      __builtin_amdgcn_readfirstlane((int)(xcd_block(blockIdx.x, gridDim.x) * (kNT / 64) + (threadIdx.x >> 6)));
  const int lane = threadIdx.x & 63;
  if (slice >= nslices) return;
  T acc = sell_row_sum<T>(S, slice, lane, sv, x);
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) acc += shfl_xor_t(acc, off, G);
  const int64_t pos = (int64_t)slice * (64 / G) + lane / G;
  if ((G == 1 || (lane & (G - 1)) == 0) && pos < n) {
    const int64_t row = S.srow ? (int64_t)S.srow[pos] : pos;
    if (const T* sd = sell_diag<T>(S)) acc += sd[row] * x[row];  // the diagonal term, last
    spmv_epilogue<MODE, T, T>(row, row, acc, y, x, b, dinv, d, xn, c1, c2);
  }
}

// k right-hand sides with k a multiple of VW = 16 / sizeof(T): L lanes per
// row, each lane owns VW consecutive columns and gathers them as one 16-byte
// load (measured for k = 16 FP32 at 161^3: 1.6x the one-column-per-lane kernel).
// Same per-column summation order as csr_multi_kernel.
template <typename T, int L, int MODE>
__global__ void __launch_bounds__(kNT) csr_multi_vec_kernel(int64_t n, int k, const int32_t* __restrict__ rp,
                                                          const int32_t* __restrict__ ci, const T* __restrict__ val,
                                                          const T* __restrict__ x, T* __restrict__ y,
                                                          const T* __restrict__ b, const T* __restrict__ dinv,
                                                          T* __restrict__ d, T* __restrict__ xn, T c1, T c2) {
  constexpr int VW = 16 / (int)sizeof(T);
  typedef T vec __attribute__((ext_vector_type(VW)));
  const int64_t row = ((int64_t)xcd_block(blockIdx.x, gridDim.x) * kNT + threadIdx.x) / L;
  const int c0 = (int)(threadIdx.x & (L - 1)) * VW;
  if (row >= n || c0 >= k) return;
  const int s = rp[row], e = rp[row + 1];
  vec acc = (vec)(T(0));
  int q = s;
  for (; q + 4 <= e; q += 4) {
    const int j0 = ci[q], j1 = ci[q + 1], j2 = ci[q + 2], j3 = ci[q + 3];
    const T v0 = val[q], v1 = val[q + 1], v2 = val[q + 2], v3 = val[q + 3];
    const vec g0 = *(const vec*)(x + (int64_t)j0 * k + c0), g1 = *(const vec*)(x + (int64_t)j1 * k + c0);
    const vec g2 = *(const vec*)(x + (int64_t)j2 * k + c0), g3 = *(const vec*)(x + (int64_t)j3 * k + c0);
    acc += v0 * g0;
    acc += v1 * g1;
    acc += v2 * g2;
    acc += v3 * g3;
  }
  for (; q < e; ++q) acc += val[q] * *(const vec*)(x + (int64_t)ci[q] * k + c0);
  const int64_t idx = row * k + c0;
  if (MODE == kAssign) {
    *(vec*)(y + idx) = acc;
  } else if (MODE == kAssignScale) {
    *(vec*)(y + idx) = acc;
    const vec v = (c1 * (T)dinv[row]) * acc;
    *(vec*)(xn + idx) = v;
    if (d) *(vec*)(d + idx) = v;
  } else if (MODE == kResid) {
    *(vec*)(y + idx) = *(const vec*)(b + idx) - acc;
  } else if (MODE == kAdd) {
    *(vec*)(y + idx) += acc;
  } else if (MODE == kJacobi) {
    *(vec*)(xn + idx) = *(const vec*)(x + idx) + (c1 * dinv[row]) * (*(const vec*)(b + idx) - acc);
  } else {  // kCheb
    const vec r = dinv[row] * (*(const vec*)(b + idx) - acc);
    const vec dn = (c1 != T(0) ? c1 * *(const vec*)(d + idx) : (vec)(T(0))) + c2 * r;
    *(vec*)(d + idx) = dn;
    *(vec*)(xn + idx) = *(const vec*)(x + idx) + dn;
  }
}

// k right-hand sides, long rows: as csr_multi_vec_kernel, and ES lane groups
// split the row's entries (group e takes entries e, e + ES, ...); a butterfly
// over the groups adds their sums (fixed order).  L * ES lanes per row.
template <typename T, int L, int ES, int MODE>
__global__ void __launch_bounds__(kNT) csr_multi_split_kernel(int64_t n, int k, const int32_t* __restrict__ rp,
                                                            const int32_t* __restrict__ ci,
                                                            const T* __restrict__ val, const T* __restrict__ x,
                                                            T* __restrict__ y, const T* __restrict__ b,
                                                            const T* __restrict__ dinv, T* __restrict__ d,
                                                            T* __restrict__ xn, T c1, T c2) {
  constexpr int VW = 16 / (int)sizeof(T);
  constexpr int LR = L * ES;
  static_assert(LR <= 64, "a row's lanes stay inside one wavefront");
  typedef T vec __attribute__((ext_vector_type(VW)));
  const int64_t row = ((int64_t)xcd_block(blockIdx.x, gridDim.x) * kNT + threadIdx.x) / LR;
  if (row >= n) return;  // all LR lanes of a row leave together
  const int sub = threadIdx.x & (LR - 1);
  const int eg = sub / L;
  const int c0 = (sub % L) * VW;
  const bool on = c0 < k;
  vec acc = (vec)(T(0));
  if (on) {
    const int s = rp[row], e = rp[row + 1];
    int q = s + eg;
    for (; q + 3 * ES < e; q += 4 * ES) {
      const int j0 = ci[q], j1 = ci[q + ES], j2 = ci[q + 2 * ES], j3 = ci[q + 3 * ES];
      const T v0 = val[q], v1 = val[q + ES], v2 = val[q + 2 * ES], v3 = val[q + 3 * ES];
      const vec g0 = *(const vec*)(x + (int64_t)j0 * k + c0), g1 = *(const vec*)(x + (int64_t)j1 * k + c0);
      const vec g2 = *(const vec*)(x + (int64_t)j2 * k + c0), g3 = *(const vec*)(x + (int64_t)j3 * k + c0);
      acc += v0 * g0;
      acc += v1 * g1;
      acc += v2 * g2;
      acc += v3 * g3;
    }
    for (; q < e; q += ES) acc += val[q] * *(const vec*)(x + (int64_t)ci[q] * k + c0);
  }
#pragma unroll
  for (int off = LR / 2; off >= L; off >>= 1) {
#pragma unroll
    for (int j = 0; j < VW; ++j) acc[j] += shfl_xor_t(acc[j], off, LR);
  }
  if (!on || eg != 0) return;
  const int64_t idx = row * k + c0;
  if (MODE == kAssign) {
    *(vec*)(y + idx) = acc;
  } else if (MODE == kAssignScale) {
    *(vec*)(y + idx) = acc;
    const vec v = (c1 * (T)dinv[row]) * acc;
    *(vec*)(xn + idx) = v;
    if (d) *(vec*)(d + idx) = v;
  } else if (MODE == kResid) {
    *(vec*)(y + idx) = *(const vec*)(b + idx) - acc;
  } else if (MODE == kAdd) {
    *(vec*)(y + idx) += acc;
  } else if (MODE == kJacobi) {
    *(vec*)(xn + idx) = *(const vec*)(x + idx) + (c1 * dinv[row]) * (*(const vec*)(b + idx) - acc);
  } else {  // kCheb
    const vec r = dinv[row] * (*(const vec*)(b + idx) - acc);
    const vec dn = (c1 != T(0) ? c1 * *(const vec*)(d + idx) : (vec)(T(0))) + c2 * r;
    *(vec*)(d + idx) = dn;
    *(vec*)(xn + idx) = *(const vec*)(x + idx) + dn;
  }
}

// k right-hand sides on the sliced-ELL copy with one lane per row in its
// layout (64-row slices): L lanes per row, each owning VW = 16 / sizeof(T)
// consecutive columns gathered as one 16-byte load; a slice takes L
// wavefronts, each over 64 / L of its rows, so a step reads the 2-byte codes
// and values of those rows as one contiguous run.  Entries in row order,
// padding (value 0) last: the per-column sums are those of
// csr_multi_vec_kernel.
template <typename T, int L, int MODE, int VW = 16 / (int)sizeof(T)>
__global__ void __launch_bounds__(kNT) sell_multi_kernel(int64_t n, int k, int64_t nslices, SellView S,
                                                       const T* __restrict__ sv, const T* __restrict__ x,
                                                       T* __restrict__ y, const T* __restrict__ b,
                                                       const T* __restrict__ dinv, T* __restrict__ d,
                                                       T* __restrict__ xn, T c1, T c2) {
  constexpr int RW = 64 / L;  // rows per wavefront
  typedef T vec __attribute__((ext_vector_type(VW)));
  const int gw = __builtin_amdgcn_readfirstlane((int)(xcd_block(blockIdx.x, gridDim.x) * (kNT / 64) + (threadIdx.x >> 6)));
  const int slice = gw / L;
  if (slice >= nslices) return;
  const int lane = threadIdx.x & 63;
  const int rl = (gw % L) * RW + lane / L;  // row position inside the slice
  const int c0 = (lane % L) * VW;
  const int off = __builtin_amdgcn_readfirstlane(S.sp[slice]);
  const int w = __builtin_amdgcn_readfirstlane((S.sp[slice + 1] - off) >> 6);
  const int32_t* wp = S.swin + slice * kSellWindows;
  const SellWin W{__builtin_amdgcn_readfirstlane(wp[0]), __builtin_amdgcn_readfirstlane(wp[1]),
                  __builtin_amdgcn_readfirstlane(wp[2]), __builtin_amdgcn_readfirstlane(wp[3]),
                  __builtin_amdgcn_readfirstlane(wp[4]), __builtin_amdgcn_readfirstlane(wp[5]),
                  __builtin_amdgcn_readfirstlane(wp[6]), __builtin_amdgcn_readfirstlane(wp[7])};
  const bool on = c0 < k;
  vec acc = (vec)(T(0));
  const T* vp = sv + off + rl;
  if (on) {
    if (W.b0 >= 0) {
      const uint16_t* cp = S.sc16 + off + rl;
      int q = 0;
      for (; q + 4 <= w; q += 4) {
        const int j0 = sell_col(cp[64 * q], W), j1 = sell_col(cp[64 * (q + 1)], W);
        const int j2 = sell_col(cp[64 * (q + 2)], W), j3 = sell_col(cp[64 * (q + 3)], W);
        const T v0 = vp[64 * q], v1 = vp[64 * (q + 1)], v2 = vp[64 * (q + 2)], v3 = vp[64 * (q + 3)];
        const vec g0 = *(const vec*)(x + (int64_t)j0 * k + c0), g1 = *(const vec*)(x + (int64_t)j1 * k + c0);
        const vec g2 = *(const vec*)(x + (int64_t)j2 * k + c0), g3 = *(const vec*)(x + (int64_t)j3 * k + c0);
        acc += v0 * g0;
        acc += v1 * g1;
        acc += v2 * g2;
        acc += v3 * g3;
      }
      for (; q < w; ++q) acc += vp[64 * q] * *(const vec*)(x + (int64_t)sell_col(cp[64 * q], W) * k + c0);
    } else {
      const int32_t* cp = S.sc + off + rl;
      int q = 0;
      for (; q + 4 <= w; q += 4) {
        const int j0 = cp[64 * q], j1 = cp[64 * (q + 1)], j2 = cp[64 * (q + 2)], j3 = cp[64 * (q + 3)];
        const T v0 = vp[64 * q], v1 = vp[64 * (q + 1)], v2 = vp[64 * (q + 2)], v3 = vp[64 * (q + 3)];
        const vec g0 = *(const vec*)(x + (int64_t)j0 * k + c0), g1 = *(const vec*)(x + (int64_t)j1 * k + c0);
        const vec g2 = *(const vec*)(x + (int64_t)j2 * k + c0), g3 = *(const vec*)(x + (int64_t)j3 * k + c0);
        acc += v0 * g0;
        acc += v1 * g1;
        acc += v2 * g2;
        acc += v3 * g3;
      }
      for (; q < w; ++q) acc += vp[64 * q] * *(const vec*)(x + (int64_t)cp[64 * q] * k + c0);
    }
  }
  const int64_t pos = (int64_t)slice * 64 + rl;
  if (!on || pos >= n) return;
  const int64_t row = S.srow ? (int64_t)S.srow[pos] : pos;
  const int64_t idx = row * k + c0;
  if (const T* sd = sell_diag<T>(S)) acc += sd[row] * *(const vec*)(x + idx);  // the diagonal term, last
  if (MODE == kAssign) {
    *(vec*)(y + idx) = acc;
  } else if (MODE == kAssignScale) {
    *(vec*)(y + idx) = acc;
    const vec v = (c1 * (T)dinv[row]) * acc;
    *(vec*)(xn + idx) = v;
    if (d) *(vec*)(d + idx) = v;
  } else if (MODE == kResid) {
    *(vec*)(y + idx) = *(const vec*)(b + idx) - acc;
  } else if (MODE == kAdd) {
    *(vec*)(y + idx) += acc;
  } else if (MODE == kJacobi) {
    *(vec*)(xn + idx) = *(const vec*)(x + idx) + (c1 * dinv[row]) * (*(const vec*)(b + idx) - acc);
  } else {  // kCheb
    const vec r = dinv[row] * (*(const vec*)(b + idx) - acc);
    const vec dn = (c1 != T(0) ? c1 * *(const vec*)(d + idx) : (vec)(T(0))) + c2 * r;
    *(vec*)(d + idx) = dn;
    *(vec*)(xn + idx) = *(const vec*)(x + idx) + dn;
  }
}

// ---- reductions ---------------------------------------------------------------

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum of v over the block (fixed order); result valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* s_buf) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) s_buf[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < kNT / 64; ++i) r += s_buf[i];
  __syncthreads();
  return r;
}

// Per-column block partials for row-major (n x k) data: thread (row, c) with
// c = tid % KP.  Writes partials[blockIdx.x * k + c] for c < k.
template <int KP>
__device__ __forceinline__ void block_col_partials(double v, int k, double* s_buf, double* partials) {
  if (KP == 1) {
    double r = block_sum(v, s_buf);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
    return;
  }
  s_buf[threadIdx.x] = v;
  __syncthreads();
  if ((int)threadIdx.x < k) {
    double s = 0.0;
    for (int j = threadIdx.x; j < kNT; j += KP) s += s_buf[j];
    partials[(int64_t)blockIdx.x * k + threadIdx.x] = s;
  }
  __syncthreads();
}

}  // namespace fem
