// V-cycle and preconditioned CG.
//
// Reference path being replaced: CG_Flex_Cycle (src/core/cuda/cgcycle.cu:7-69),
// AMG_Level::cycle / cycle_level0 (src/core/cuda/amg_level.cu:22-128) and the
// smoother kernels of src/core/cuda/gauss_seidel.cu (preRRRFullSymmetric
// :2018-2261, postPCRFullSymmetric :4409-4662).
//
// Differences by design (results are compared on the solution, which the
// preconditioner does not change):
//  * the V-cycle is symmetric (same polynomial smoother before and after the
//    coarse correction), so plain CG is valid for it;
//  * the outer residual stays FP64 and is never overwritten by the V-cycle;
//  * the stop test is re-checked on the true residual b - A x;
//  * scalars (alpha, beta, norms) stay on the device; the host reads one
//    vector of squared residual norms per iteration.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>

#include "fem_internal.hpp"
#include "kernels.hpp"

namespace fem {

namespace {

enum { S_RZ = 0, S_PQ, S_ALPHA, S_BETA, S_RR, S_BB, S_ACT, S_RZN, S_TRUE, S_SEQ, S_COUNT };

// host-mapped record of one |r|^2 finish: act[64], rr[64], sequence number;
// two records (by the parity of the sequence number)
constexpr int kActRec = 136;

int pow2_at_least(int k) {
  int p = 1;
  while (p < k) p <<= 1;
  return p;
}

// lanes per row of the CSR kernels, from the average row length (measured at
// 161^3: 64-entry coarse rows run fastest with 16 lanes, 100-entry
// restriction rows with 32)
int lanes_for(const Csr& A) {
  double avg = A.n ? (double)A.nnz / (double)A.n : 1.0;
  if (avg <= 6) return 4;
  if (avg <= 12) return 8;
  if (avg <= 80) return 16;
  if (avg <= 200) return 32;
  return 64;
}

const Csr& level_A(const Handle& h, int l) { return l == 0 ? h.A : h.H.levels[l]->A; }

template <typename V>
const V* vals_of(const Csr& A);
template <>
const double* vals_of<double>(const Csr& A) { return A.v64.get(); }
template <>
const float* vals_of<float>(const Csr& A) { return A.v32.get(); }

template <typename V>
const V* dinv_of(const Level& L);
template <>
const double* dinv_of<double>(const Level& L) { return L.dinv64.get(); }
template <>
const float* dinv_of<float>(const Level& L) { return L.dinv32.get(); }

template <typename T>
struct LV;
template <>
struct LV<double> {
  static double* x(Level& L) { return L.x64.get(); }
  static double* t(Level& L) { return L.t64.get(); }
  static double* b(Level& L) { return L.b64.get(); }
  static double* r(Level& L) { return L.r64.get(); }
  static double* d(Level& L) { return L.d64.get(); }
};
template <>
struct LV<float> {
  static float* x(Level& L) { return L.x32.get(); }
  static float* t(Level& L) { return L.t32.get(); }
  static float* b(Level& L) { return L.b32.get(); }
  static float* r(Level& L) { return L.r32.get(); }
  static float* d(Level& L) { return L.d32.get(); }
};

template <typename V, typename T, int MODE, int G>
void launch_vec(const Csr& A, const V* val, const T* x, T* y, const T* b, const V* dinv, T* d, T* xn, T c1, T c2,
                hipStream_t s) {
  unsigned grid = grid_for(A.n * G, kNT);
  csr_vec_kernel<V, T, G, MODE><<<grid, kNT, 0, s>>>(A.n, A.rp.get(), A.ci.get(), val, x, y, b, dinv, d, xn, c1, c2);
}

template <typename T, int MODE, int L, int ES>
void launch_multi_split(const Csr& A, int k, const T* val, const T* x, T* y, const T* b, const T* dinv, T* d, T* xn,
                        T c1, T c2, hipStream_t s) {
  unsigned grid = grid_for(A.n * L * ES, kNT);
  csr_multi_split_kernel<T, L, ES, MODE><<<grid, kNT, 0, s>>>(A.n, k, A.rp.get(), A.ci.get(), val, x, y, b, dinv, d,
                                                              xn, c1, c2);
}

// Entry groups per row of the multi-column CSR kernels: enough for at most
// ~16 entries each (coarse operators, restrictions), within one wavefront.
int entry_split(const Csr& A, int L) {
  const double avg = A.n ? (double)A.nnz / (double)A.n : 0.0;
  int es = 1;
  while (es * L < 64 && avg > 16.0 * es) es *= 2;
  return es;
}

template <typename T, int MODE, int L>
void launch_multi_vec(const Csr& A, int k, const T* val, const T* x, T* y, const T* b, const T* dinv, T* d, T* xn,
                      T c1, T c2, hipStream_t s) {
  switch (L < 64 ? entry_split(A, L) : 1) {
    case 1: break;
    case 2: launch_multi_split<T, MODE, L, (L <= 32 ? 2 : 1)>(A, k, val, x, y, b, dinv, d, xn, c1, c2, s); return;
    case 4: launch_multi_split<T, MODE, L, (L <= 16 ? 4 : 1)>(A, k, val, x, y, b, dinv, d, xn, c1, c2, s); return;
    case 8: launch_multi_split<T, MODE, L, (L <= 8 ? 8 : 1)>(A, k, val, x, y, b, dinv, d, xn, c1, c2, s); return;
    case 16: launch_multi_split<T, MODE, L, (L <= 4 ? 16 : 1)>(A, k, val, x, y, b, dinv, d, xn, c1, c2, s); return;
    case 32: launch_multi_split<T, MODE, L, (L <= 2 ? 32 : 1)>(A, k, val, x, y, b, dinv, d, xn, c1, c2, s); return;
    default: launch_multi_split<T, MODE, L, (L <= 1 ? 64 : 1)>(A, k, val, x, y, b, dinv, d, xn, c1, c2, s); return;
  }
  unsigned grid = grid_for(A.n * L, kNT);
  csr_multi_vec_kernel<T, L, MODE><<<grid, kNT, 0, s>>>(A.n, k, A.rp.get(), A.ci.get(), val, x, y, b, dinv, d, xn,
                                                        c1, c2);
}

template <typename V, typename T, int MODE, int KP>
void launch_multi(const Csr& A, int k, const V* val, const T* x, T* y, const T* b, const V* dinv, T* d, T* xn, T c1,
                  T c2, hipStream_t s) {
  unsigned grid = grid_for(A.n * KP, kNT);
  csr_multi_kernel<V, T, KP, MODE><<<grid, kNT, 0, s>>>(A.n, k, A.rp.get(), A.ci.get(), val, x, y, b, dinv, d, xn,
                                                       c1, c2);
}

template <typename T>
const T* svals_of(const Csr& A);
template <>
const double* svals_of<double>(const Csr& A) { return A.s64.get(); }
template <>
const float* svals_of<float>(const Csr& A) { return A.s32.get(); }

template <typename T, int MODE, int G>
void launch_sell(const Csr& A, const T* x, T* y, const T* b, const T* dinv, T* d, T* xn, T c1, T c2, hipStream_t s) {
  sell_kernel<T, MODE, G><<<grid_for(A.nslices, kNT / 64), kNT, 0, s>>>(A.n, A.nslices, sell_view(A), svals_of<T>(A), x,
                                                                       y, b, dinv, d, xn, c1, c2);
}

// columns per lane of the multi-column sliced kernel: 4 (FP32 16-byte, FP64
// 32-byte gathers; measured at 161^3 with 16 columns: FP64 with 2 per lane
// 12 % slower, FP32 with 8 per lane 7 % slower)
template <typename T>
constexpr int sell_multi_cols() { return 4; }
// longer average rows take the entry-split CSR kernel for k columns (161^3
// level 1, ~45 entries per row: 45 us against ~69 us on the sliced copy)
constexpr int64_t kMultiSellRow = 24;

template <typename T, int MODE, int L>
void launch_sell_multi(const Csr& A, int k, const T* x, T* y, const T* b, const T* dinv, T* d, T* xn, T c1, T c2,
                       hipStream_t s) {
  unsigned grid = grid_for(A.nslices * L, kNT / 64);
  sell_multi_kernel<T, L, MODE, sell_multi_cols<T>()><<<grid, kNT, 0, s>>>(A.n, k, A.nslices, sell_view(A),
                                                                           svals_of<T>(A), x, y, b, dinv, d, xn, c1, c2);
}

// y/xn = f(A x) for k right-hand sides; MODE selects the epilogue.  One RHS
// uses the sliced-ELL copy when it exists, otherwise CSR with G lanes per row.
template <typename T, int MODE>
void spmv_op(const Csr& A, int k, const T* x, T* y, const T* b, const T* dinv, T* d, T* xn, T c1, T c2,
             hipStream_t s) {
  if (A.n == 0) return;
  const T* val = vals_of<T>(A);
  if (k == 1 && A.nslices > 0 && svals_of<T>(A) && A.lds_bs > 0) {
    const unsigned nb = (unsigned)((A.nslices + A.lds_bs - 1) / A.lds_bs);
    sell_lds_kernel<T, MODE><<<nb, kNT, sizeof(T) * (size_t)A.lds_maxc, s>>>(
        A.n, A.nslices, sell_view(A), lds_view(A), svals_of<T>(A), x, y, b, dinv, d, xn, c1, c2);
  } else if (k == 1 && A.nslices > 0 && svals_of<T>(A)) {
    switch (A.sell_g) {
      case 1: launch_sell<T, MODE, 1>(A, x, y, b, dinv, d, xn, c1, c2, s); break;
      case 2: launch_sell<T, MODE, 2>(A, x, y, b, dinv, d, xn, c1, c2, s); break;
      case 4: launch_sell<T, MODE, 4>(A, x, y, b, dinv, d, xn, c1, c2, s); break;
      case 8: launch_sell<T, MODE, 8>(A, x, y, b, dinv, d, xn, c1, c2, s); break;
      case 16: launch_sell<T, MODE, 16>(A, x, y, b, dinv, d, xn, c1, c2, s); break;
      default: launch_sell<T, MODE, 32>(A, x, y, b, dinv, d, xn, c1, c2, s); break;
    }
  } else if (k == 1) {
    switch (lanes_for(A)) {
      case 4: launch_vec<T, T, MODE, 4>(A, val, x, y, b, dinv, d, xn, c1, c2, s); break;
      case 8: launch_vec<T, T, MODE, 8>(A, val, x, y, b, dinv, d, xn, c1, c2, s); break;
      case 16: launch_vec<T, T, MODE, 16>(A, val, x, y, b, dinv, d, xn, c1, c2, s); break;
      case 32: launch_vec<T, T, MODE, 32>(A, val, x, y, b, dinv, d, xn, c1, c2, s); break;
      default: launch_vec<T, T, MODE, 64>(A, val, x, y, b, dinv, d, xn, c1, c2, s); break;
    }
  } else if (k % sell_multi_cols<T>() == 0 && A.nslices > 0 && A.sell_g == 1 && svals_of<T>(A) &&
             A.nnz <= kMultiSellRow * A.n) {
    // column groups on the sliced-ELL copy
    switch (pow2_at_least(k / sell_multi_cols<T>())) {
      case 1: launch_sell_multi<T, MODE, 1>(A, k, x, y, b, dinv, d, xn, c1, c2, s); break;
      case 2: launch_sell_multi<T, MODE, 2>(A, k, x, y, b, dinv, d, xn, c1, c2, s); break;
      case 4: launch_sell_multi<T, MODE, 4>(A, k, x, y, b, dinv, d, xn, c1, c2, s); break;
      case 8: launch_sell_multi<T, MODE, 8>(A, k, x, y, b, dinv, d, xn, c1, c2, s); break;
      case 16: launch_sell_multi<T, MODE, 16>(A, k, x, y, b, dinv, d, xn, c1, c2, s); break;
      default: launch_sell_multi<T, MODE, 32>(A, k, x, y, b, dinv, d, xn, c1, c2, s); break;
    }
  } else if (k % (16 / (int)sizeof(T)) == 0) {
    // 16-byte column groups (all multi-column vectors here come from hipMalloc)
    switch (pow2_at_least(k / (16 / (int)sizeof(T)))) {
      case 1: launch_multi_vec<T, MODE, 1>(A, k, val, x, y, b, dinv, d, xn, c1, c2, s); break;
      case 2: launch_multi_vec<T, MODE, 2>(A, k, val, x, y, b, dinv, d, xn, c1, c2, s); break;
      case 4: launch_multi_vec<T, MODE, 4>(A, k, val, x, y, b, dinv, d, xn, c1, c2, s); break;
      case 8: launch_multi_vec<T, MODE, 8>(A, k, val, x, y, b, dinv, d, xn, c1, c2, s); break;
      case 16: launch_multi_vec<T, MODE, 16>(A, k, val, x, y, b, dinv, d, xn, c1, c2, s); break;
      default: launch_multi_vec<T, MODE, 32>(A, k, val, x, y, b, dinv, d, xn, c1, c2, s); break;
    }
  } else {
    switch (pow2_at_least(k)) {
      case 2: launch_multi<T, T, MODE, 2>(A, k, val, x, y, b, dinv, d, xn, c1, c2, s); break;
      case 4: launch_multi<T, T, MODE, 4>(A, k, val, x, y, b, dinv, d, xn, c1, c2, s); break;
      case 8: launch_multi<T, T, MODE, 8>(A, k, val, x, y, b, dinv, d, xn, c1, c2, s); break;
      case 16: launch_multi<T, T, MODE, 16>(A, k, val, x, y, b, dinv, d, xn, c1, c2, s); break;
      case 32: launch_multi<T, T, MODE, 32>(A, k, val, x, y, b, dinv, d, xn, c1, c2, s); break;
      default: launch_multi<T, T, MODE, 64>(A, k, val, x, y, b, dinv, d, xn, c1, c2, s); break;
    }
  }
  HIP_LAUNCH_CHECK();
}

// ---- elementwise kernels ------------------------------------------------------

// x = c1 * dinv * b (first smoothing step from a zero guess); d = x if d != null
template <typename V, typename Tb, typename T>
__global__ void scale_dinv_kernel(int64_t n, int k, const V* __restrict__ dinv, const Tb* __restrict__ b,
                                  T* __restrict__ x, T* __restrict__ d, T c1) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * k) return;
  int64_t i = t / k;
  T v = c1 * (T)dinv[i] * (T)b[t];
  x[t] = v;
  if (d) d[t] = v;
}

// First level-0 smoothing step of a mixed-precision V-cycle, from the FP64
// Krylov residual: b = fl32(r); x = c1 * dinv * b; d = x if d != null.
__global__ void scale_in_kernel(int64_t n, const float* __restrict__ dinv, const double* __restrict__ r,
                                float* __restrict__ b, float* __restrict__ x, float* __restrict__ d, float c1) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float bv = (float)r[i];
  b[i] = bv;
  const float v = c1 * dinv[i] * bv;
  x[i] = v;
  if (d) d[i] = v;
}

// Last level-0 smoothing step of a mixed-precision V-cycle (one column,
// sliced ELL): the smoothed value goes straight to the FP64 output z, with
// per-block partials of r.z for the PCG beta (d is not needed afterwards).
template <int MODE>
__global__ void __launch_bounds__(kNT) sell_smooth_out_kernel(int64_t n, int64_t nslices, SellView S,
                                                            const float* __restrict__ sv, const float* __restrict__ x,
                                                            const float* __restrict__ b, const float* __restrict__ dinv,
                                                            const float* __restrict__ d, float c1, float c2,
                                                            const double* __restrict__ r, double* __restrict__ z,
                                                            double* __restrict__ partials) {
  __shared__ double s_buf[kNT];
  const int slice =
      __builtin_amdgcn_readfirstlane((int)(xcd_block(blockIdx.x, gridDim.x) * (kNT / 64) + (threadIdx.x >> 6)));
  const int lane = threadIdx.x & 63;
  double part = 0.0;
  if (slice < nslices) {
    float acc = sell_row_sum<float>(S, slice, lane, sv, x);
    const int64_t row = (int64_t)slice * 64 + lane;
    if (row < n) {
      if (S.sd32) acc += S.sd32[row] * x[row];  // the diagonal term, last
      float xn;
      if (MODE == kJacobi) {
        xn = x[row] + c1 * dinv[row] * (b[row] - acc);
      } else {
        const float rr = ntload(dinv + row) * (ntload(b + row) - acc);
        const float dn = (c1 != 0.0f ? c1 * ntload(d + row) : 0.0f) + c2 * rr;
        xn = x[row] + dn;
      }
      const double zv = (double)xn;
      z[row] = zv;
      part = r[row] * zv;
    }
  }
  const double tot = block_sum(part, s_buf);
  if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

// The same on the LDS-staged copy: a block per L.bs slices, r.z partials per block.
template <int MODE>
__global__ void __launch_bounds__(kNT) sell_lds_out_kernel(int64_t n, int64_t nslices, SellView S, LdsView L,
                                                         const float* __restrict__ sv, const float* __restrict__ x,
                                                         const float* __restrict__ b, const float* __restrict__ dinv,
                                                         const float* __restrict__ d, float c1, float c2,
                                                         const double* __restrict__ r, double* __restrict__ z,
                                                         double* __restrict__ partials) {
  extern __shared__ unsigned char lds_raw[];
  float* xs = reinterpret_cast<float*>(lds_raw);
  __shared__ double s_buf[kNT];
  const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  lds_stage<float>(L, blk, x, xs);
  const int lane = threadIdx.x & 63;
  double part = 0.0;
  for (int q = threadIdx.x >> 6; q < L.bs; q += kNT / 64) {
    const int slice = __builtin_amdgcn_readfirstlane((int)(blk * L.bs + q));
    if (slice >= nslices) break;
    float acc = lds_row_sum<float>(S, L, slice, lane, sv, xs);
    const int64_t row = (int64_t)slice * 64 + lane;
    if (row < n) {
      if (S.sd32) acc += S.sd32[row] * x[row];  // the diagonal term, last
      float xn;
      if (MODE == kJacobi) {
        xn = x[row] + c1 * dinv[row] * (b[row] - acc);
      } else {
        const float rr = ntload(dinv + row) * (ntload(b + row) - acc);
        const float dn = (c1 != 0.0f ? c1 * ntload(d + row) : 0.0f) + c2 * rr;
        xn = x[row] + dn;
      }
      const double zv = (double)xn;
      z[row] = zv;
      part += r[row] * zv;
    }
  }
  const double tot = block_sum(part, s_buf);
  if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

// ---- precision 2: level-0 smoothing in y = D^1/2 x on D^-1/2 A D^-1/2 with
// FP16 off-diagonals (diagonal 1 + eps_i held apart in FP32)

// first step from y = 0: b^ = s fl32(r), y = c1 b^ (d = y if d != null)
__global__ void scale_in_half_kernel(int64_t n, const float* __restrict__ sd, const double* __restrict__ r,
                                     float* __restrict__ bh, float* __restrict__ y, float* __restrict__ d, float c1) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float bv = sd[i] * (float)r[i];
  bh[i] = bv;
  const float v = c1 * bv;
  y[i] = v;
  if (d) d[i] = v;
}

// r^ = b^ - (diag_i y_i + sum_j h_ij y_j); MODE kResid: out = r^;
// kJacobi: yn = y + c1 r^; kCheb: d = c1 d + c2 r^, yn = y + d
template <int MODE>
__global__ void __launch_bounds__(kNT) sell_half_kernel(int64_t n, int64_t nslices, SellView S,
                                                      const uint16_t* __restrict__ sh, const float* __restrict__ diag,
                                                      const float* __restrict__ y, const float* __restrict__ bh,
                                                      float* __restrict__ out, float* __restrict__ d,
                                                      float* __restrict__ yn, float c1, float c2) {
  const int slice =
      __builtin_amdgcn_readfirstlane((int)(xcd_block(blockIdx.x, gridDim.x) * (kNT / 64) + (threadIdx.x >> 6)));
  const int lane = threadIdx.x & 63;
  if (slice >= nslices) return;
  const float acc = sell_row_sum<float, uint16_t>(S, slice, lane, sh, y);
  const int64_t row = (int64_t)slice * 64 + lane;
  if (row >= n) return;
  const float yi = y[row];
  const float r = ntload(bh + row) - (ntload(diag + row) * yi + acc);
  if (MODE == kResid) {
    out[row] = r;
  } else if (MODE == kJacobi) {
    yn[row] = yi + c1 * r;
  } else {
    const float dn = (c1 != 0.0f ? c1 * ntload(d + row) : 0.0f) + c2 * r;
    d[row] = dn;
    yn[row] = yi + dn;
  }
}

// last post-smoothing step: z = s y_new in FP64 and per-block partials of r.z
template <int MODE>
__global__ void __launch_bounds__(kNT) sell_half_out_kernel(int64_t n, int64_t nslices, SellView S,
                                                          const uint16_t* __restrict__ sh,
                                                          const float* __restrict__ diag, const float* __restrict__ sd,
                                                          const float* __restrict__ y, const float* __restrict__ bh,
                                                          const float* __restrict__ d, float c1, float c2,
                                                          const double* __restrict__ r64, double* __restrict__ z,
                                                          double* __restrict__ partials) {
  __shared__ double s_buf[kNT];
  const int slice =
      __builtin_amdgcn_readfirstlane((int)(xcd_block(blockIdx.x, gridDim.x) * (kNT / 64) + (threadIdx.x >> 6)));
  const int lane = threadIdx.x & 63;
  double part = 0.0;
  if (slice < nslices) {
    const float acc = sell_row_sum<float, uint16_t>(S, slice, lane, sh, y);
    const int64_t row = (int64_t)slice * 64 + lane;
    if (row < n) {
      const float yi = y[row];
      const float r = ntload(bh + row) - (ntload(diag + row) * yi + acc);
      float yv;
      if (MODE == kJacobi) {
        yv = yi + c1 * r;
      } else {
        const float dn = (c1 != 0.0f ? c1 * ntload(d + row) : 0.0f) + c2 * r;
        yv = yi + dn;
      }
      const double zv = (double)(sd[row] * yv);
      z[row] = zv;
      part = r64[row] * zv;
    }
  }
  const double tot = block_sum(part, s_buf);
  if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

// The same two kernels on the LDS-staged copy: a block stages y at its
// slices' distinct columns once, and every entry reads y from LDS through its
// 16-bit index (same entries in the same order, so the same sums).
template <int MODE>
__global__ void __launch_bounds__(kNT) sell_half_lds_kernel(int64_t n, int64_t nslices, SellView S, LdsView L,
                                                          const uint16_t* __restrict__ sh,
                                                          const float* __restrict__ diag, const float* __restrict__ y,
                                                          const float* __restrict__ bh, float* __restrict__ out,
                                                          float* __restrict__ d, float* __restrict__ yn, float c1,
                                                          float c2) {
  extern __shared__ unsigned char lds_raw[];
  float* ys = reinterpret_cast<float*>(lds_raw);
  const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  lds_stage<float>(L, blk, y, ys);
  const int lane = threadIdx.x & 63;
  for (int q = threadIdx.x >> 6; q < L.bs; q += kNT / 64) {
    const int slice = __builtin_amdgcn_readfirstlane((int)(blk * L.bs + q));
    if (slice >= nslices) break;
    const float acc = lds_row_sum<float, uint16_t>(S, L, slice, lane, sh, ys);
    const int64_t row = (int64_t)slice * 64 + lane;
    if (row >= n) continue;
    const float yi = y[row];
    const float r = ntload(bh + row) - (ntload(diag + row) * yi + acc);
    if (MODE == kResid) {
      out[row] = r;
    } else if (MODE == kJacobi) {
      yn[row] = yi + c1 * r;
    } else {
      const float dn = (c1 != 0.0f ? c1 * ntload(d + row) : 0.0f) + c2 * r;
      d[row] = dn;
      yn[row] = yi + dn;
    }
  }
}

template <int MODE>
__global__ void __launch_bounds__(kNT) sell_half_lds_out_kernel(int64_t n, int64_t nslices, SellView S, LdsView L,
                                                              const uint16_t* __restrict__ sh,
                                                              const float* __restrict__ diag,
                                                              const float* __restrict__ sd,
                                                              const float* __restrict__ y,
                                                              const float* __restrict__ bh,
                                                              const float* __restrict__ d, float c1, float c2,
                                                              const double* __restrict__ r64, double* __restrict__ z,
                                                              double* __restrict__ partials) {
  extern __shared__ unsigned char lds_raw[];
  float* ys = reinterpret_cast<float*>(lds_raw);
  __shared__ double s_buf[kNT];
  const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  lds_stage<float>(L, blk, y, ys);
  const int lane = threadIdx.x & 63;
  double part = 0.0;
  for (int q = threadIdx.x >> 6; q < L.bs; q += kNT / 64) {
    const int slice = __builtin_amdgcn_readfirstlane((int)(blk * L.bs + q));
    if (slice >= nslices) break;
    const float acc = lds_row_sum<float, uint16_t>(S, L, slice, lane, sh, ys);
    const int64_t row = (int64_t)slice * 64 + lane;
    if (row >= n) continue;
    const float yi = y[row];
    const float r = ntload(bh + row) - (ntload(diag + row) * yi + acc);
    float yv;
    if (MODE == kJacobi) {
      yv = yi + c1 * r;
    } else {
      const float dn = (c1 != 0.0f ? c1 * ntload(d + row) : 0.0f) + c2 * r;
      yv = yi + dn;
    }
    const double zv = (double)(sd[row] * yv);
    z[row] = zv;
    part += r64[row] * zv;
  }
  const double tot = block_sum(part, s_buf);
  if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

template <typename Tin, typename Tout>
__global__ void convert_kernel(int64_t m, const Tin* __restrict__ a, Tout* __restrict__ b) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < m) b[t] = (Tout)a[t];
}

// dense coarse solve x = C b with the FP64 inverse C (row-major nc x nc) in
// every precision: b is widened to FP64, the products are summed in FP64 and
// only the result is rounded to the level's vector type; one wave per row
template <typename T>
__global__ void __launch_bounds__(64) coarse_kernel(int64_t nc, int k, const double* __restrict__ C,
                                                   const T* __restrict__ b, T* __restrict__ x) {
  const int64_t i = blockIdx.x;
  const int lane = threadIdx.x;
  for (int c = 0; c < k; ++c) {
    double acc = 0.0;
    for (int64_t j = lane; j < nc; j += 64) acc += C[i * nc + j] * (double)b[j * k + c];
    acc = wave_sum(acc);
    if (lane == 0) x[i * k + c] = (T)acc;
  }
}

// ---- PCG vector kernels (FP64, k columns, row-major) --------------------------
// Each reduction kernel writes one partial per block and column; a one-block
// finish kernel sums them in a fixed order and updates the device scalars.
// (A last-block-done pattern inside the producer would need an agent-scope
// release fence per block, which writes back the XCD's L2 on gfx950.)

// q = A p and per-block partials of p.q (one column)
template <int G>
__global__ void __launch_bounds__(kNT) pcg_spmv_dot_kernel(int64_t n, const int32_t* __restrict__ rp,
                                                         const int32_t* __restrict__ ci, const double* __restrict__ val,
                                                         const double* __restrict__ p, double* __restrict__ q,
                                                         double* __restrict__ partials) {
  __shared__ double s_buf[kNT];
  double part = 0.0;
  const int lane = threadIdx.x & (G - 1);
  // kNT / G rows per pass, one front per XCD
  const XcdRows R = xcd_rows(n, kNT / G);
  for (int64_t row = R.first + threadIdx.x / G; row < R.hi; row += R.step) {
    const int s = rp[row], e = rp[row + 1];
    double acc = 0.0;
    for (int j = s + lane; j < e; j += G) acc += ntload(val + j) * p[ntload(ci + j)];
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, G);
    if (lane == 0) {
      q[row] = acc;
      part += acc * p[row];
    }
  }
  double tot = block_sum(part, s_buf);
  if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

// q = A p on the sliced-ELL copy and per-block partials of p.q (one column);
// a wavefront per slice as in the smoother kernels (no grid-stride loop).
__global__ void __launch_bounds__(kNT) pcg_sell_dot_kernel(int64_t n, int64_t nslices, SellView S,
                                                         const double* __restrict__ sv, const double* __restrict__ p,
                                                         double* __restrict__ q, double* __restrict__ partials) {
  __shared__ double s_buf[kNT];
  const int lane = threadIdx.x & 63;
  const int slice =
      __builtin_amdgcn_readfirstlane((int)(xcd_block(blockIdx.x, gridDim.x) * (kNT / 64) + (threadIdx.x >> 6)));
  double part = 0.0;
  if (slice < nslices) {
    double acc = sell_row_sum<double>(S, slice, lane, sv, p);
    const int64_t row = (int64_t)slice * 64 + lane;
    if (row < n) {
      if (S.sd64) acc += S.sd64[row] * p[row];  // the diagonal term, last
      q[row] = acc;
      part = acc * p[row];
    }
  }
  double tot = block_sum(part, s_buf);
  if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

// The same on the LDS-staged copy: a block per L.bs slices, p.q partials per block.
__global__ void __launch_bounds__(kNT) pcg_sell_lds_dot_kernel(int64_t n, int64_t nslices, SellView S, LdsView L,
                                                             const double* __restrict__ sv,
                                                             const double* __restrict__ p, double* __restrict__ q,
                                                             double* __restrict__ partials) {
  extern __shared__ unsigned char lds_raw[];
  double* xs = reinterpret_cast<double*>(lds_raw);
  __shared__ double s_buf[kNT];
  const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  lds_stage<double>(L, blk, p, xs);
  const int lane = threadIdx.x & 63;
  double part = 0.0;
  for (int w = threadIdx.x >> 6; w < L.bs; w += kNT / 64) {
    const int slice = __builtin_amdgcn_readfirstlane((int)(blk * L.bs + w));
    if (slice >= nslices) break;
    double acc = lds_row_sum<double>(S, L, slice, lane, sv, xs);
    const int64_t row = (int64_t)slice * 64 + lane;
    if (row < n) {
      if (S.sd64) acc += S.sd64[row] * p[row];  // the diagonal term, last
      q[row] = acc;
      part += acc * p[row];
    }
  }
  double tot = block_sum(part, s_buf);
  if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

// q = A p and per-block, per-column partials of p.q (k columns)
template <int KP>
__global__ void __launch_bounds__(kNT) pcg_spmm_dot_kernel(int64_t n, int k, const int32_t* __restrict__ rp,
                                                         const int32_t* __restrict__ ci,
                                                         const double* __restrict__ val, const double* __restrict__ p,
                                                         double* __restrict__ q, double* __restrict__ partials) {
  __shared__ double s_buf[kNT];
  double part = 0.0;
  const int c = threadIdx.x & (KP - 1);
  // kNT / KP rows per pass, one front per XCD
  const XcdRows R = xcd_rows(n, kNT / KP);
  if (c < k) {
    for (int64_t row = R.first + threadIdx.x / KP; row < R.hi; row += R.step) {
      const int s = rp[row], e = rp[row + 1];
      double acc = 0.0;
      int j = s;
      for (; j + 4 <= e; j += 4) {
        const double g0 = p[(int64_t)ci[j] * k + c], g1 = p[(int64_t)ci[j + 1] * k + c],
                     g2 = p[(int64_t)ci[j + 2] * k + c], g3 = p[(int64_t)ci[j + 3] * k + c];
        acc += val[j] * g0;
        acc += val[j + 1] * g1;
        acc += val[j + 2] * g2;
        acc += val[j + 3] * g3;
      }
      for (; j < e; ++j) acc += val[j] * p[(int64_t)ci[j] * k + c];
      q[row * k + c] = acc;
      part += acc * p[row * k + c];
    }
  }
  block_col_partials<KP>(part, k, s_buf, partials);
}

// x += alpha p; r -= alpha q; partials of r.r
// (and r32 = fl32(r), the FP32 V-cycle's input, when r32 is given)
template <int KP>
__global__ void __launch_bounds__(kNT) pcg_update_kernel(int64_t n, int k, double* __restrict__ x,
                                                       double* __restrict__ r, const double* __restrict__ p,
                                                       const double* __restrict__ q, double* __restrict__ partials,
                                                       const double* __restrict__ scal, float* __restrict__ r32,
                                                       const float* __restrict__ dinv32, float* __restrict__ x0,
                                                       float* __restrict__ d0, float c0) {
  __shared__ double s_buf[kNT];
  double part = 0.0;
  const int c = threadIdx.x & (KP - 1);
  const int64_t rows_per_grid = (int64_t)gridDim.x * (kNT / KP);
  if (c < k) {
    const double a = scal[S_ALPHA * k + c];
    for (int64_t row = ((int64_t)blockIdx.x * kNT + threadIdx.x) / KP; row < n; row += rows_per_grid) {
      int64_t t = row * k + c;
      x[t] += a * p[t];
      double rv = r[t] - a * q[t];
      r[t] = rv;
      if (r32) {
        const float bv = (float)rv;
        r32[t] = bv;
        if (x0) {  // the level-0 smoother's first step from zero on fl32(r)
          const float v = c0 * dinv32[row] * bv;
          x0[t] = v;
          if (d0) d0[t] = v;
        }
      }
      part += rv * rv;
    }
  }
  block_col_partials<KP>(part, k, s_buf, partials);
}

// The same for an even k: L lanes per row, two columns per lane as one
// 16-byte load; per-column sums in the same order as pcg_spmm_dot_kernel.
template <int L>
__global__ void __launch_bounds__(kNT) pcg_spmm_dot_vec_kernel(int64_t n, int k, const int32_t* __restrict__ rp,
                                                             const int32_t* __restrict__ ci,
                                                             const double* __restrict__ val,
                                                             const double* __restrict__ p, double* __restrict__ q,
                                                             double* __restrict__ partials) {
  typedef double vec __attribute__((ext_vector_type(2)));
  __shared__ double s_buf[2 * kNT];
  const int c0 = (int)(threadIdx.x & (L - 1)) * 2;
  const XcdRows R = xcd_rows(n, kNT / L);
  vec part = (vec)(0.0);
  if (c0 < k) {
    for (int64_t row = R.first + threadIdx.x / L; row < R.hi; row += R.step) {
      const int s = rp[row], e = rp[row + 1];
      vec acc = (vec)(0.0);
      int j = s;
      for (; j + 4 <= e; j += 4) {
        const vec g0 = *(const vec*)(p + (int64_t)ci[j] * k + c0), g1 = *(const vec*)(p + (int64_t)ci[j + 1] * k + c0);
        const vec g2 = *(const vec*)(p + (int64_t)ci[j + 2] * k + c0),
                  g3 = *(const vec*)(p + (int64_t)ci[j + 3] * k + c0);
        acc += val[j] * g0;
        acc += val[j + 1] * g1;
        acc += val[j + 2] * g2;
        acc += val[j + 3] * g3;
      }
      for (; j < e; ++j) acc += val[j] * *(const vec*)(p + (int64_t)ci[j] * k + c0);
      *(vec*)(q + row * k + c0) = acc;
      part += acc * *(const vec*)(p + row * k + c0);
    }
  }
  s_buf[2 * threadIdx.x] = part.x;
  s_buf[2 * threadIdx.x + 1] = part.y;
  __syncthreads();
  if ((int)threadIdx.x < k) {
    const int c = threadIdx.x;
    double sum = 0.0;
    for (int t = c / 2; t < kNT; t += L) sum += s_buf[2 * t + (c & 1)];
    partials[(int64_t)blockIdx.x * k + c] = sum;
  }
}

// q = A p on the sliced-ELL copy (one lane per row in its layout) for an even
// k, with per-block, per-column partials of p.q.  As sell_multi_kernel, L
// lanes per row own two columns each; a unit is one slice quarter (64 / L
// rows), walked by the block's wavefronts one front per XCD.
template <int L, int CW>
__global__ void __launch_bounds__(kNT) pcg_sell_multi_dot_kernel(int64_t n, int k, int64_t nslices, SellView S,
                                                                const double* __restrict__ sv,
                                                                const double* __restrict__ p, double* __restrict__ q,
                                                                double* __restrict__ partials) {
  typedef double vec __attribute__((ext_vector_type(CW)));
  constexpr int RW = 64 / L;
  __shared__ double s_buf[CW * kNT];
  const int lane = threadIdx.x & 63;
  const int c0 = (lane % L) * CW;
  const bool on = c0 < k;
  const XcdRows R = xcd_rows(nslices * L, kNT / 64);
  vec part = (vec)(0.0);
  for (int64_t u = R.first + (threadIdx.x >> 6); u < R.hi; u += R.step) {
    const int unit = __builtin_amdgcn_readfirstlane((int)u);
    const int slice = unit / L;
    const int rl = (unit % L) * RW + lane / L;
    const int off = __builtin_amdgcn_readfirstlane(S.sp[slice]);
    const int w = __builtin_amdgcn_readfirstlane((S.sp[slice + 1] - off) >> 6);
    const int32_t* wp = S.swin + slice * kSellWindows;
    const SellWin W{__builtin_amdgcn_readfirstlane(wp[0]), __builtin_amdgcn_readfirstlane(wp[1]),
                    __builtin_amdgcn_readfirstlane(wp[2]), __builtin_amdgcn_readfirstlane(wp[3]),
                    __builtin_amdgcn_readfirstlane(wp[4]), __builtin_amdgcn_readfirstlane(wp[5]),
                    __builtin_amdgcn_readfirstlane(wp[6]), __builtin_amdgcn_readfirstlane(wp[7])};
    const int64_t pos = (int64_t)slice * 64 + rl;
    if (!on || pos >= n) continue;
    const double* vp = sv + off + rl;
    vec acc = (vec)(0.0);
    if (W.b0 >= 0) {
      const uint16_t* cp = S.sc16 + off + rl;
      int j = 0;
      for (; j + 4 <= w; j += 4) {
        const int j0 = sell_col(cp[64 * j], W), j1 = sell_col(cp[64 * (j + 1)], W);
        const int j2 = sell_col(cp[64 * (j + 2)], W), j3 = sell_col(cp[64 * (j + 3)], W);
        const double v0 = vp[64 * j], v1 = vp[64 * (j + 1)], v2 = vp[64 * (j + 2)], v3 = vp[64 * (j + 3)];
        const vec g0 = *(const vec*)(p + (int64_t)j0 * k + c0), g1 = *(const vec*)(p + (int64_t)j1 * k + c0);
        const vec g2 = *(const vec*)(p + (int64_t)j2 * k + c0), g3 = *(const vec*)(p + (int64_t)j3 * k + c0);
        acc += v0 * g0;
        acc += v1 * g1;
        acc += v2 * g2;
        acc += v3 * g3;
      }
      for (; j < w; ++j) acc += vp[64 * j] * *(const vec*)(p + (int64_t)sell_col(cp[64 * j], W) * k + c0);
    } else {
      const int32_t* cp = S.sc + off + rl;
      for (int j = 0; j < w; ++j) acc += vp[64 * j] * *(const vec*)(p + (int64_t)cp[64 * j] * k + c0);
    }
    if (S.sd64) acc += S.sd64[pos] * *(const vec*)(p + pos * k + c0);  // the diagonal term, last
    *(vec*)(q + pos * k + c0) = acc;
    part += acc * *(const vec*)(p + pos * k + c0);
  }
#pragma unroll
  for (int j = 0; j < CW; ++j) s_buf[CW * threadIdx.x + j] = part[j];
  __syncthreads();
  if ((int)threadIdx.x < k) {
    const int c = threadIdx.x;
    double sum = 0.0;
    for (int t = c / CW; t < kNT; t += L) sum += s_buf[CW * t + (c % CW)];
    partials[(int64_t)blockIdx.x * k + c] = sum;
  }
}

// Last level-0 smoothing step for k columns (a multiple of 4) on the FP32
// sliced copy, with per-block, per-column partials of r.z (r the FP64 Krylov
// residual, z the step's FP32 output, which is written; d is not, the step
// being the last).  Lanes and units as in pcg_sell_multi_dot_kernel (L lanes
// per row, 4 columns each, one sweep front per XCD); per column z is what
// sell_multi_kernel computes for the same step.
template <int MODE, int L>
__global__ void __launch_bounds__(kNT) sell_multi_out_kernel(int64_t n, int k, int64_t nslices, SellView S,
                                                           const float* __restrict__ sv, const float* __restrict__ x,
                                                           const float* __restrict__ b,
                                                           const float* __restrict__ dinv,
                                                           const float* __restrict__ d, float* __restrict__ xn,
                                                           float c1, float c2, const double* __restrict__ r,
                                                           double* __restrict__ partials) {
  typedef float vec __attribute__((ext_vector_type(4)));
  constexpr int RW = 64 / L;
  __shared__ double s_buf[4 * kNT];
  const int lane = threadIdx.x & 63;
  const int c0 = (lane % L) * 4;
  const bool on = c0 < k;
  const XcdRows R = xcd_rows(nslices * L, kNT / 64);
  double p0 = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0;
  for (int64_t u = R.first + (threadIdx.x >> 6); u < R.hi; u += R.step) {
    const int unit = __builtin_amdgcn_readfirstlane((int)u);
    const int slice = unit / L;
    const int rl = (unit % L) * RW + lane / L;
    const int off = __builtin_amdgcn_readfirstlane(S.sp[slice]);
    const int w = __builtin_amdgcn_readfirstlane((S.sp[slice + 1] - off) >> 6);
    const int32_t* wp = S.swin + slice * kSellWindows;
    const SellWin W{__builtin_amdgcn_readfirstlane(wp[0]), __builtin_amdgcn_readfirstlane(wp[1]),
                    __builtin_amdgcn_readfirstlane(wp[2]), __builtin_amdgcn_readfirstlane(wp[3]),
                    __builtin_amdgcn_readfirstlane(wp[4]), __builtin_amdgcn_readfirstlane(wp[5]),
                    __builtin_amdgcn_readfirstlane(wp[6]), __builtin_amdgcn_readfirstlane(wp[7])};
    const int64_t pos = (int64_t)slice * 64 + rl;
    if (!on || pos >= n) continue;
    const float* vp = sv + off + rl;
    vec acc = (vec)(0.0f);
    if (W.b0 >= 0) {
      const uint16_t* cp = S.sc16 + off + rl;
      int q = 0;
      for (; q + 4 <= w; q += 4) {
        const int j0 = sell_col(cp[64 * q], W), j1 = sell_col(cp[64 * (q + 1)], W);
        const int j2 = sell_col(cp[64 * (q + 2)], W), j3 = sell_col(cp[64 * (q + 3)], W);
        const float v0 = vp[64 * q], v1 = vp[64 * (q + 1)], v2 = vp[64 * (q + 2)], v3 = vp[64 * (q + 3)];
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
        const float v0 = vp[64 * q], v1 = vp[64 * (q + 1)], v2 = vp[64 * (q + 2)], v3 = vp[64 * (q + 3)];
        const vec g0 = *(const vec*)(x + (int64_t)j0 * k + c0), g1 = *(const vec*)(x + (int64_t)j1 * k + c0);
        const vec g2 = *(const vec*)(x + (int64_t)j2 * k + c0), g3 = *(const vec*)(x + (int64_t)j3 * k + c0);
        acc += v0 * g0;
        acc += v1 * g1;
        acc += v2 * g2;
        acc += v3 * g3;
      }
      for (; q < w; ++q) acc += vp[64 * q] * *(const vec*)(x + (int64_t)cp[64 * q] * k + c0);
    }
    const int64_t idx = pos * k + c0;
    if (S.sd32) acc += S.sd32[pos] * *(const vec*)(x + idx);  // the diagonal term, last
    vec z;
    if (MODE == kJacobi) {
      z = *(const vec*)(x + idx) + (c1 * dinv[pos]) * (*(const vec*)(b + idx) - acc);
    } else {  // kCheb
      const vec rr = dinv[pos] * (*(const vec*)(b + idx) - acc);
      const vec dn = (c1 != 0.0f ? c1 * *(const vec*)(d + idx) : (vec)(0.0f)) + c2 * rr;
      z = *(const vec*)(x + idx) + dn;
    }
    *(vec*)(xn + idx) = z;
    p0 += r[idx] * (double)z[0];
    p1 += r[idx + 1] * (double)z[1];
    p2 += r[idx + 2] * (double)z[2];
    p3 += r[idx + 3] * (double)z[3];
  }
  s_buf[4 * threadIdx.x] = p0;
  s_buf[4 * threadIdx.x + 1] = p1;
  s_buf[4 * threadIdx.x + 2] = p2;
  s_buf[4 * threadIdx.x + 3] = p3;
  __syncthreads();
  if ((int)threadIdx.x < k) {
    const int c = threadIdx.x;
    double sum = 0.0;
    for (int t = c / 4; t < kNT; t += L) sum += s_buf[4 * t + (c % 4)];
    partials[(int64_t)blockIdx.x * k + c] = sum;
  }
}

// partials of a.b per column
template <int KP, typename B>
__global__ void __launch_bounds__(kNT) pcg_dot_kernel(int64_t n, int k, const double* __restrict__ a,
                                                    const B* __restrict__ b, double* __restrict__ partials) {
  __shared__ double s_buf[kNT];
  double part = 0.0;
  const int c = threadIdx.x & (KP - 1);
  const int64_t rows_per_grid = (int64_t)gridDim.x * (kNT / KP);
  if (c < k)
    for (int64_t row = ((int64_t)blockIdx.x * kNT + threadIdx.x) / KP; row < n; row += rows_per_grid)
      part += a[row * k + c] * (double)b[row * k + c];
  block_col_partials<KP>(part, k, s_buf, partials);
}

enum FinishMode { F_STORE = 0, F_ALPHA = 1, F_RR = 2, F_BETA = 3, F_FLEX = 4 };

// one block: column totals of nb partials, then the scalar update
// 1024 threads so that ~16k partials take a few rounds of independent loads.
constexpr int kFinNT = 1024;

__global__ void __launch_bounds__(kFinNT) pcg_finish_kernel(int nb, int k, const double* __restrict__ partials,
                                                          double* __restrict__ scal, int mode, int slot, double tol2,
                                                          double* __restrict__ host_act) {
  __shared__ double s_buf[kFinNT / 64];
  __shared__ double s_seq;
  // |r|^2 finishes are numbered on the device; record n goes to host record n % 2
  if (mode == F_RR && threadIdx.x == 0) {
    s_seq = scal[S_SEQ * k] + 1.0;
    scal[S_SEQ * k] = s_seq;
  }
  __syncthreads();
  double* rec = (mode == F_RR && host_act) ? host_act + ((int64_t)s_seq & 1) * kActRec : nullptr;
  for (int c = 0; c < k; ++c) {
    double v = 0.0;
#pragma unroll 8
    for (int b = threadIdx.x; b < nb; b += kFinNT) v += partials[(int64_t)b * k + c];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) s_buf[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      v = 0.0;
      for (int w = 0; w < kFinNT / 64; ++w) v += s_buf[w];
      if (mode == F_STORE) {
        scal[slot * k + c] = v;
      } else if (mode == F_ALPHA) {
        scal[S_PQ * k + c] = v;
        scal[S_ALPHA * k + c] = (scal[S_ACT * k + c] != 0.0 && v != 0.0) ? scal[S_RZ * k + c] / v : 0.0;
      } else if (mode == F_RR) {
        scal[S_RR * k + c] = v;
        bool act = scal[S_ACT * k + c] != 0.0 && v > tol2 * scal[S_BB * k + c];
        scal[S_ACT * k + c] = act ? 1.0 : 0.0;
        if (rec) {
          rec[c] = act ? 1.0 : 0.0;
          rec[64 + c] = v;
        }
      } else if (mode == F_FLEX) {
        // v = z.q with q = A p of the last step: z.r_old = z.r + alpha z.q, so
        // the flexible (Polak-Ribiere) beta = (z.r - z.r_old) / rz_old =
        // -alpha z.q / rz_old; z.r was stored in S_RZN
        const double rz = scal[S_RZ * k + c];
        scal[S_BETA * k + c] = (scal[S_ACT * k + c] != 0.0 && rz != 0.0) ? -scal[S_ALPHA * k + c] * v / rz : 0.0;
        scal[S_RZ * k + c] = scal[S_RZN * k + c];
      } else {
        double rz = scal[S_RZ * k + c];
        scal[S_BETA * k + c] = (scal[S_ACT * k + c] != 0.0 && rz != 0.0) ? v / rz : 0.0;
        scal[S_RZ * k + c] = v;
      }
    }
    __syncthreads();
  }
  if (rec && threadIdx.x == 0) rec[128] = s_seq;  // read by the host after this kernel completed
}

// p = z + beta p
// p = z + beta p (z as the preconditioner left it: FP64, or the FP32 V-cycle
// result, converted exactly)
template <typename Z>
__global__ void pcg_p_kernel(int64_t m, int k, const Z* __restrict__ z, double* __restrict__ p,
                             const double* __restrict__ scal) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m) return;
  int c = (int)(t % k);
  p[t] = (double)z[t] + scal[S_BETA * k + c] * p[t];
}

// x[:, c] = 0 for every column c whose bit is set in mask (k <= 64 columns)
__global__ void zero_cols_kernel(int64_t m, int k, unsigned long long mask, double* __restrict__ x) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < m && ((mask >> (t % k)) & 1ull)) x[t] = 0.0;
}

__global__ void copy_kernel(int64_t m, const double* __restrict__ a, double* __restrict__ b) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < m) b[t] = a[t];
}

unsigned vec_grid(int64_t n, int kp) {
  int64_t g = (n * kp + kNT - 1) / kNT;
  if (g > 1024) g = 1024;
  if (g < 1) g = 1;
  return (unsigned)g;
}

// ---- V-cycle ---------------------------------------------------------------------

struct Smoother {
  int type;        // 0 Jacobi, 1 Chebyshev
  double omega;    // Jacobi weight
  double lmin, lmax;
};

Smoother smoother_for(const Handle& h, const Level& L) {
  Smoother S;
  S.type = h.prm.smootherType;
  S.omega = h.prm.smootherWeight * (4.0 / 3.0) / L.rho;
  S.lmax = 1.1 * L.rho;
  S.lmin = S.lmax / 30.0;
  return S;
}

// Bytes of the stored matrix one SpMV with k columns reads: a value and a
// column index per nonzero; the sliced-ELL copy (one column) stores 2-byte
// column codes in its compact slices.  Padding is not counted.
template <typename T>
double matrix_bytes(const Csr& A, int k) {
  const bool sell = k == 1 && A.nslices > 0;
  const double nd = sell ? (double)A.nnz_sdiag : 0.0;  // diagonal outside the slices: no column code
  if (sell && A.lds_bs > 0)  // 16-bit LDS indices, and each block's column list
    return (double)A.nnz * sizeof(T) + 2.0 * ((double)A.nnz - nd) + 4.0 * (double)A.lcols.size();
  const double nc = sell ? (double)A.nnz_compact : 0.0;
  return (double)A.nnz * sizeof(T) + 2.0 * nc + 4.0 * ((double)A.nnz - nd - nc);
}

template <typename T>
void fill_zero(T* p, int64_t n, hipStream_t s);
template <>
void fill_zero<double>(double* p, int64_t n, hipStream_t s) { dev_fill_f64(p, 0.0, (size_t)n, s); }
template <>
void fill_zero<float>(float* p, int64_t n, hipStream_t s) { dev_fill_f32(p, 0.0f, (size_t)n, s); }

// nu steps of the smoother from x = 0 (zero == true) or from the given x.
// Buffers x and t ping-pong; returns the buffer that holds the result.
// Fused FP64 ends of a mixed-precision level-0 smoother (one column, sliced
// ELL): the first step reads the Krylov residual r64 (and stores fl32(r64) in
// b32 for the later steps), the last step writes z64 and r64.z64 partials.
struct Level0Io {
  const double* r64 = nullptr;
  float* b32 = nullptr;
  double* z64 = nullptr;
  double* partials = nullptr;
  unsigned nb = 0;  // partials written by the last step
};

template <typename T>
bool fused_out(const Level0Io* io, const Csr& A, int k) {
  return std::is_same<T, float>::value && io && io->z64 && k == 1 && A.nslices > 0 && A.s32.size() > 0 &&
         A.srow.size() == 0 && A.sell_g == 1;
}

// The last level-0 smoothing step for k > 1 columns with r.z partials
// (io->r64 and io->partials, no io->z64) on the FP32 sliced copy
template <typename T>
bool multi_out(const Level0Io* io, const Csr& A, int k) {
  return std::is_same<T, float>::value && io && io->r64 && io->partials && !io->z64 && k > 1 && k % 4 == 0 &&
         k <= 64 && A.nslices > 0 && A.s32.size() > 0 && A.srow.size() == 0 && A.sell_g == 1;
}

// launches sell_multi_out_kernel; returns the partial blocks written
template <int MODE>
unsigned launch_multi_out(const Csr& A, int k, const float* x, const float* b, const float* dinv, const float* d,
                          float* xn, float c1, float c2, const double* r, double* partials, hipStream_t s) {
  const int L = pow2_at_least(k / 4);
  int64_t g = (A.nslices * L + kNT / 64 - 1) / (kNT / 64);
  if (g > 4096) g = 4096;
  const SellView S = sell_view(A);
  const float* sv = A.s32.get();
  switch (L) {
    case 1:
      sell_multi_out_kernel<MODE, 1><<<(unsigned)g, kNT, 0, s>>>(A.n, k, A.nslices, S, sv, x, b, dinv, d, xn, c1, c2,
                                                                 r, partials);
      break;
    case 2:
      sell_multi_out_kernel<MODE, 2><<<(unsigned)g, kNT, 0, s>>>(A.n, k, A.nslices, S, sv, x, b, dinv, d, xn, c1, c2,
                                                                 r, partials);
      break;
    case 4:
      sell_multi_out_kernel<MODE, 4><<<(unsigned)g, kNT, 0, s>>>(A.n, k, A.nslices, S, sv, x, b, dinv, d, xn, c1, c2,
                                                                 r, partials);
      break;
    case 8:
      sell_multi_out_kernel<MODE, 8><<<(unsigned)g, kNT, 0, s>>>(A.n, k, A.nslices, S, sv, x, b, dinv, d, xn, c1, c2,
                                                                 r, partials);
      break;
    default:
      sell_multi_out_kernel<MODE, 16><<<(unsigned)g, kNT, 0, s>>>(A.n, k, A.nslices, S, sv, x, b, dinv, d, xn, c1,
                                                                  c2, r, partials);
      break;
  }
  HIP_LAUNCH_CHECK();
  return (unsigned)g;
}

// nu steps of the smoother from x = 0 (zero == true) or from the given x.
// Buffers x and t ping-pong; returns the buffer that holds the result (with
// io->z64 set, the last step writes there instead and nullptr is returned;
// with multi_out, the last step also writes the r.z partials).
// first_done: the first step from zero is already in x (and d), taken by the
// restriction that produced b.
template <typename T>
T* smooth(Handle& h, int l, const Csr& A, const T* b, T* x, T* t, T* d, int nu, bool zero, int k,
          Level0Io* io = nullptr, bool first_done = false) {
  Level& L = *h.H.levels[l];
  hipStream_t s = h.stream;
  const T* dinv = dinv_of<T>(L);
  Smoother S = smoother_for(h, L);
  const bool prof = h.profiling && l == 0;
  // algorithmic bytes of one sweep: the matrix and the per-row vectors
  // (Jacobi: x, b, dinv, x_new; Chebyshev: + d read and write)
  const double sweep_bytes = matrix_bytes<T>(A, k) + (double)A.n * sizeof(T) * (S.type ? 6.0 : 4.0) * k;
  // fused last step: x, b, dinv (+ d) in FP32, r read and z written in FP64
  const double out_bytes = matrix_bytes<T>(A, k) + (double)A.n * (4.0 * (S.type ? 4.0 : 3.0) + 16.0);
  if (nu <= 0) {
    // a fill kernel, not a memset: this may run inside a captured graph
    if (zero) fill_zero<T>(x, A.n * k, s);
    return x;
  }
  const bool in64 = zero && io && io->r64 && std::is_same<T, float>::value && k == 1;
  const bool out64 = !zero && fused_out<T>(io, A, k);
  const bool outk = !zero && multi_out<T>(io, A, k);
  auto first_step = [&](T* xo, T* dout, double c) {
    if (in64) {
      scale_in_kernel<<<grid_for(A.n, 256), 256, 0, s>>>(A.n, (const float*)dinv, io->r64, io->b32, (float*)xo,
                                                          (float*)dout, (float)c);
    } else {
      scale_dinv_kernel<T, T, T><<<grid_for(A.n * k, 256), 256, 0, s>>>(A.n, k, dinv, b, xo, dout, (T)c);
    }
    HIP_LAUNCH_CHECK();
  };
  auto last_step = [&](int mode, T* cur, double c1, double c2) {
    if (A.lds_bs > 0) {
      const unsigned nb = (unsigned)((A.nslices + A.lds_bs - 1) / A.lds_bs);
      const size_t lds = sizeof(float) * (size_t)A.lds_maxc;
      if (mode == kJacobi)
        sell_lds_out_kernel<kJacobi><<<nb, kNT, lds, s>>>(A.n, A.nslices, sell_view(A), lds_view(A), A.s32.get(),
                                                          (const float*)cur, (const float*)b, (const float*)dinv,
                                                          (const float*)d, (float)c1, (float)c2, io->r64, io->z64,
                                                          io->partials);
      else
        sell_lds_out_kernel<kCheb><<<nb, kNT, lds, s>>>(A.n, A.nslices, sell_view(A), lds_view(A), A.s32.get(),
                                                        (const float*)cur, (const float*)b, (const float*)dinv,
                                                        (const float*)d, (float)c1, (float)c2, io->r64, io->z64,
                                                        io->partials);
      HIP_LAUNCH_CHECK();
      io->nb = nb;
      return;
    }
    const unsigned g = grid_for(A.nslices, kNT / 64);
    if (mode == kJacobi)
      sell_smooth_out_kernel<kJacobi><<<g, kNT, 0, s>>>(A.n, A.nslices, sell_view(A), A.s32.get(), (const float*)cur,
                                                         (const float*)b, (const float*)dinv, (const float*)d,
                                                         (float)c1, (float)c2, io->r64, io->z64, io->partials);
    else
      sell_smooth_out_kernel<kCheb><<<g, kNT, 0, s>>>(A.n, A.nslices, sell_view(A), A.s32.get(), (const float*)cur,
                                                       (const float*)b, (const float*)dinv, (const float*)d, (float)c1,
                                                       (float)c2, io->r64, io->z64, io->partials);
    HIP_LAUNCH_CHECK();
    io->nb = g;
  };
  T* cur = x;
  T* nxt = t;
  int step = 0;
  if (S.type == 0) {
    const T w = (T)S.omega;
    if (zero) {
      if (!first_done) first_step(cur, nullptr, S.omega);
      step = 1;
    }
    for (; step < nu; ++step) {
      if (out64 && step == nu - 1) {
        ProfScope ps(h, prof ? "smooth_l0_out" : "", out_bytes);
        last_step(kJacobi, cur, S.omega, 0.0);
        return nullptr;
      }
      if constexpr (std::is_same<T, float>::value) {
        if (outk && step == nu - 1) {
          io->nb = launch_multi_out<kJacobi>(A, k, cur, b, dinv, nullptr, nxt, (float)S.omega, 0.0f, io->r64,
                                             io->partials, s);
          return nxt;
        }
      }
      ProfScope ps(h, prof ? "smooth_l0" : "", sweep_bytes);
      spmv_op<T, kJacobi>(A, k, cur, nullptr, b, dinv, nullptr, nxt, w, T(0), s);
      T* tmp = cur; cur = nxt; nxt = tmp;
    }
    return cur;
  }
  // Chebyshev on D^-1 A over [lmin, lmax]
  const double theta = 0.5 * (S.lmax + S.lmin), delta = 0.5 * (S.lmax - S.lmin);
  const double sigma = theta / delta;
  double rho_old = 1.0 / sigma;
  if (zero) {
    if (!first_done) first_step(cur, d, 1.0 / theta);
    step = 1;
  }
  for (; step < nu; ++step) {
    double c1, c2;
    if (step == 0) {
      c1 = 0.0;
      c2 = 1.0 / theta;
    } else {
      double rho_new = 1.0 / (2.0 * sigma - rho_old);
      c1 = rho_new * rho_old;
      c2 = 2.0 * rho_new / delta;
      rho_old = rho_new;
    }
    if (out64 && step == nu - 1) {
      ProfScope ps(h, prof ? "smooth_l0_out" : "", out_bytes);
      last_step(kCheb, cur, c1, c2);
      return nullptr;
    }
    if constexpr (std::is_same<T, float>::value) {
      if (outk && step == nu - 1) {
        io->nb = launch_multi_out<kCheb>(A, k, cur, b, dinv, d, nxt, (float)c1, (float)c2, io->r64, io->partials, s);
        return nxt;
      }
    }
    ProfScope ps(h, prof ? "smooth_l0" : "", sweep_bytes);
    spmv_op<T, kCheb>(A, k, cur, nullptr, b, dinv, d, nxt, (T)c1, (T)c2, s);
    T* tmp = cur; cur = nxt; nxt = tmp;
  }
  return cur;
}

// b_{l+1} = R r.  When level l+1 is smoothed from zero (not the coarsest
// level) the same pass takes its first smoothing step, x = c D^-1 b (and
// d = x for Chebyshev), as smooth() would; returns whether it did.
template <typename T>
bool restrict_first(Handle& h, int l, const Csr& R, const T* r, int k) {
  Hierarchy& H = h.H;
  Level& N = *H.levels[l + 1];
  T* bc = LV<T>::b(N);
  if (l + 1 == H.coarse || h.prm.preInnerIters < 1) {
    spmv_op<T, kAssign>(R, k, r, bc, nullptr, (const T*)nullptr, nullptr, nullptr, T(0), T(0), h.stream);
    return false;
  }
  const Smoother S = smoother_for(h, N);
  const double c = S.type ? 1.0 / (0.5 * (S.lmax + S.lmin)) : S.omega;
  spmv_op<T, kAssignScale>(R, k, r, bc, nullptr, dinv_of<T>(N), S.type ? LV<T>::d(N) : nullptr, LV<T>::x(N), (T)c,
                           T(0), h.stream);
  return true;
}

// Cycle types (cycleType): 0 V; 1 W, whose coarse correction is two W-cycles
// of the next level, the second from the first's result (cycles/wcycle.h);
// 2 F, whose coarse correction is an F-cycle of the next level followed by a
// V-cycle from its result (the textbook F-cycle; cycles/fcycle.h writes a
// W-cycle first).  3 (K) is a V-cycle: the reference dispatches K_CYCLE as
// the outer CG_Flex_Cycle itself (cycles/cycle.h:34-36, amg.cu:160-165).
constexpr int kCycleV = 0, kCycleW = 1, kCycleF = 2;

int cycle_type(const Handle& h) {
  const int c = h.prm.cycleType;
  return (c == kCycleW || c == kCycleF) ? c : kCycleV;
}

template <typename T>
T* vcycle(Handle& h, int l, const T* b, int k, Level0Io* io = nullptr, bool first_done = false, int type = kCycleV,
          T* x0 = nullptr);

// Coarse correction of level l, the right-hand side already restricted into
// level l+1's b: one cycle of level l+1 (V), or two (W: W then W; F: F then
// V).  Returns level l+1's buffer holding the correction.
template <typename T>
T* coarse_correction(Handle& h, int l, int k, bool next_first, int type) {
  Level& N = *h.H.levels[l + 1];
  T* xc = vcycle<T>(h, l + 1, LV<T>::b(N), k, nullptr, next_first, type);
  if (type != kCycleV && l + 1 != h.H.coarse)
    xc = vcycle<T>(h, l + 1, LV<T>::b(N), k, nullptr, false, type == kCycleW ? kCycleW : kCycleV, xc);
  return xc;
}

// One cycle of the given type on level l for the right-hand side b (level l
// vectors, k columns), from zero or (x0, one of level l's x / t buffers) from
// x0; returns the buffer holding the result (nullptr when io->z64 received
// it).  first_done: restrict_first took the first pre-smoothing step.
template <typename T>
T* vcycle(Handle& h, int l, const T* b, int k, Level0Io* io, bool first_done, int type, T* x0) {
  Hierarchy& H = h.H;
  Level& L = *H.levels[l];
  hipStream_t s = h.stream;
  const Csr& A = level_A(h, l);
  if (l == H.coarse) {  // the direct solve does not depend on a start
    coarse_kernel<T><<<(unsigned)H.nc, 64, 0, s>>>(H.nc, k, H.cinv64.get(), b, LV<T>::x(L));
    HIP_LAUNCH_CHECK();
    return LV<T>::x(L);
  }
  if (x0 && (io || first_done)) fail(FEM_ERR_STATE, "a cycle from a given start takes no fused ends");
  T* x = x0 ? x0 : LV<T>::x(L);
  T* t = (x == LV<T>::x(L)) ? LV<T>::t(L) : LV<T>::x(L);
  T* r = LV<T>::r(L);
  T* d = LV<T>::d(L);
  const int nu_pre = h.prm.preInnerIters;
  const int nu_post = h.prm.postInnerIters * (h.prm.postRelaxes > 0 ? h.prm.postRelaxes : 1);
  T* cur = smooth<T>(h, l, A, b, x, t, d, nu_pre, x0 == nullptr, k, x0 ? nullptr : io, first_done);
  T* oth = (cur == x) ? t : x;
  {
    ProfScope ps(h, (h.profiling && l == 0) ? "resid_l0" : "", matrix_bytes<T>(A, k) + 3.0 * A.n * k * sizeof(T));
    spmv_op<T, kResid>(A, k, cur, r, b, (const T*)nullptr, nullptr, nullptr, T(0), T(0), s);
  }
  const bool next_first = restrict_first<T>(h, l, L.R, r, k);
  T* xc = coarse_correction<T>(h, l, k, next_first, type);
  spmv_op<T, kAdd>(L.P, k, xc, cur, nullptr, (const T*)nullptr, nullptr, nullptr, T(0), T(0), s);
  cur = smooth<T>(h, l, A, b, cur, oth, d, nu_post, false, k, io);
  return cur;
}

// Chebyshev coefficients of step `step` of a smoother on [lmin, lmax]
struct ChebSeq {
  double theta, delta, sigma, rho_old;
  explicit ChebSeq(const Smoother& S)
      : theta(0.5 * (S.lmax + S.lmin)), delta(0.5 * (S.lmax - S.lmin)), sigma(theta / delta), rho_old(delta / theta) {}
  void next(int step, double& c1, double& c2) {
    if (step == 0) {
      c1 = 0.0;
      c2 = 1.0 / theta;
    } else {
      const double rho_new = 1.0 / (2.0 * sigma - rho_old);
      c1 = rho_new * rho_old;
      c2 = 2.0 * rho_new / delta;
      rho_old = rho_new;
    }
  }
};

// Precision 2, one column: the level-0 part of the V-cycle in y = D^1/2 x (the
// same V-cycle as the FP32 one: Chebyshev/Jacobi on a matrix similar to
// D^-1 A, restriction by (D^1/2 P)^T, prolongation by D^1/2 P), with the
// FP64 residual in and z = D^-1/2 y out.  Returns the count of r.z partials.
unsigned vcycle0_half(Handle& h, const double* r64, double* z64, double* partials) {
  Hierarchy& H = h.H;
  Level& L0 = *H.levels[0];
  hipStream_t s = h.stream;
  const Csr& O = H.ah;
  const int64_t n = O.n;
  float* bh = L0.b32.get();
  float* rr = L0.r32.get();
  float* d = L0.d32.get();
  float* cur = L0.x32.get();
  float* nxt = L0.t32.get();
  const Smoother S = smoother_for(h, L0);
  const int nu_pre = h.prm.preInnerIters;
  const int nu_post = h.prm.postInnerIters * (h.prm.postRelaxes > 0 ? h.prm.postRelaxes : 1);
  const unsigned g = grid_for(O.nslices, kNT / 64);
  const SellView V = sell_view(O);
  // LDS-staged copy when its lists exist: one block per O.lds_bs slices
  const bool staged = O.lds_bs > 0;
  const LdsView LV_ = staged ? lds_view(O) : LdsView{};
  const unsigned gl = staged ? (unsigned)((O.nslices + O.lds_bs - 1) / O.lds_bs) : 0;
  const size_t ldsb = staged ? sizeof(float) * (size_t)O.lds_maxc : 0;
  // algorithmic bytes: FP16 value + column code per off-diagonal, and y, b^, diag (+ d r/w), y_new
  const double mbytes = staged ? 2.0 * (double)O.nnz + 2.0 * (double)O.nnz + 4.0 * (double)O.lcols.size()
                               : 2.0 * (double)O.nnz + 2.0 * (double)O.nnz_compact +
                                     4.0 * ((double)O.nnz - O.nnz_compact);
  const double step_bytes = mbytes + 4.0 * n * (S.type ? 6.0 : 4.0);
  const bool prof = h.profiling;
  auto step = [&](int mode, double c1, double c2) {
    ProfScope ps(h, prof ? "smooth_l0_h" : "", step_bytes);
    if (staged && mode == kJacobi)
      sell_half_lds_kernel<kJacobi><<<gl, kNT, ldsb, s>>>(n, O.nslices, V, LV_, O.sh.get(), H.ah_diag.get(), cur,
                                                           bh, nullptr, nullptr, nxt, (float)c1, 0.0f);
    else if (staged)
      sell_half_lds_kernel<kCheb><<<gl, kNT, ldsb, s>>>(n, O.nslices, V, LV_, O.sh.get(), H.ah_diag.get(), cur, bh,
                                                         nullptr, d, nxt, (float)c1, (float)c2);
    else if (mode == kJacobi)
      sell_half_kernel<kJacobi><<<g, kNT, 0, s>>>(n, O.nslices, V, O.sh.get(), H.ah_diag.get(), cur, bh, nullptr,
                                                  nullptr, nxt, (float)c1, 0.0f);
    else
      sell_half_kernel<kCheb><<<g, kNT, 0, s>>>(n, O.nslices, V, O.sh.get(), H.ah_diag.get(), cur, bh, nullptr, d,
                                                nxt, (float)c1, (float)c2);
    HIP_LAUNCH_CHECK();
    std::swap(cur, nxt);
  };
  // pre-smoothing from y = 0
  ChebSeq cs(S);
  scale_in_half_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, H.ah_s.get(), r64, bh, cur, S.type ? d : nullptr,
                                                         (float)(S.type ? 1.0 / cs.theta : S.omega));
  HIP_LAUNCH_CHECK();
  for (int st = 1; st < nu_pre; ++st) {
    double c1 = S.omega, c2 = 0.0;
    if (S.type) cs.next(st, c1, c2);
    step(S.type ? kCheb : kJacobi, c1, c2);
  }
  // residual, restriction, coarse correction, prolongation
  if (staged)
    sell_half_lds_kernel<kResid><<<gl, kNT, ldsb, s>>>(n, O.nslices, V, LV_, O.sh.get(), H.ah_diag.get(), cur, bh,
                                                        rr, nullptr, nullptr, 0.0f, 0.0f);
  else
    sell_half_kernel<kResid><<<g, kNT, 0, s>>>(n, O.nslices, V, O.sh.get(), H.ah_diag.get(), cur, bh, rr, nullptr,
                                                nullptr, 0.0f, 0.0f);
  HIP_LAUNCH_CHECK();
  const bool next_first = restrict_first<float>(h, 0, H.rh, rr, 1);
  float* xc = coarse_correction<float>(h, 0, 1, next_first, cycle_type(h));
  spmv_op<float, kAdd>(H.ph, 1, xc, cur, nullptr, (const float*)nullptr, nullptr, nullptr, 0.0f, 0.0f, s);
  // post-smoothing; the last step writes z and the r.z partials
  ChebSeq cp(S);
  for (int st = 0; st < nu_post; ++st) {
    double c1 = S.omega, c2 = 0.0;
    if (S.type) cp.next(st, c1, c2);
    if (st < nu_post - 1) {
      step(S.type ? kCheb : kJacobi, c1, c2);
      continue;
    }
    ProfScope ps(h, prof ? "smooth_l0_h_out" : "", step_bytes + 12.0 * n);
    if (staged) {
      if (S.type)
        sell_half_lds_out_kernel<kCheb><<<gl, kNT, ldsb, s>>>(n, O.nslices, V, LV_, O.sh.get(), H.ah_diag.get(),
                                                               H.ah_s.get(), cur, bh, d, (float)c1, (float)c2, r64,
                                                               z64, partials);
      else
        sell_half_lds_out_kernel<kJacobi><<<gl, kNT, ldsb, s>>>(n, O.nslices, V, LV_, O.sh.get(), H.ah_diag.get(),
                                                                 H.ah_s.get(), cur, bh, nullptr, (float)c1, 0.0f,
                                                                 r64, z64, partials);
      HIP_LAUNCH_CHECK();
      continue;
    }
    if (S.type)
      sell_half_out_kernel<kCheb><<<g, kNT, 0, s>>>(n, O.nslices, V, O.sh.get(), H.ah_diag.get(), H.ah_s.get(), cur,
                                                    bh, d, (float)c1, (float)c2, r64, z64, partials);
    else
      sell_half_out_kernel<kJacobi><<<g, kNT, 0, s>>>(n, O.nslices, V, O.sh.get(), H.ah_diag.get(), H.ah_s.get(),
                                                      cur, bh, nullptr, (float)c1, 0.0f, r64, z64, partials);
    HIP_LAUNCH_CHECK();
  }
  return staged ? gl : g;
}

// Where one application of the preconditioner left z = M r.
struct PrecOut {
  const double* z64 = nullptr;  // FP64 z: the caller's buffer or a level-0 work vector
  const float* z32 = nullptr;   // or the FP32 V-cycle result (exact in FP64)
  unsigned nb = 0;              // r.z partials written by the fused level-0 ends
};

// One column, mixed precision, level 0 smoothed before and after the coarse
// correction, sliced ELL: the FP64 <-> FP32 conversions are fused into the
// first and last level-0 smoothing steps, and the last one also writes
// per-block partials of r.z.
struct First0 {
  bool on = false;
  float* b32 = nullptr;
  const float* dinv = nullptr;
  float* x0 = nullptr;
  float* d0 = nullptr;
  float c0 = 0.0f;
};

bool fused_level0(const Handle& h, int k, bool partials) {
  const int nu_post = h.prm.postInnerIters * (h.prm.postRelaxes > 0 ? h.prm.postRelaxes : 1);
  return partials && h.H.precision >= 1 && k == 1 && h.H.coarse > 0 && h.prm.preInnerIters >= 1 && nu_post >= 1 &&
         h.A.nslices > 0 && h.A.s32.size() > 0 && h.A.sell_g == 1;
}

// Outputs of the PCG update for an FP32 level-0 V-cycle: b32 = fl32(r), and
// the smoother's first step from zero, x0 = c0 dinv b32 (and d0 = x0 for
// Chebyshev).  Off for the FP16 level-0 smoother (its own first step) and for
// hierarchies without level-0 smoothing.
First0 first0_for(Handle& h, int k) {
  First0 f;
  if (h.H.precision < 1 || h.H.coarse <= 0 || h.prm.preInnerIters < 1) return f;
  if (h.H.precision == 2 && fused_level0(h, k, true) && h.H.ah.nslices > 0) return f;
  Level& L0 = *h.H.levels[0];
  const Smoother S = smoother_for(h, L0);
  f.b32 = L0.b32.get();
  f.dinv = L0.dinv32.get();
  f.x0 = L0.x32.get();
  if (S.type == 0) {
    f.c0 = (float)S.omega;
  } else {
    f.c0 = (float)(1.0 / (0.5 * (S.lmax + S.lmin)));
    f.d0 = L0.d32.get();
  }
  f.on = f.dinv != nullptr && f.x0 != nullptr;
  return f;
}

// z = M r at level 0 (FP64 r); k columns.  z is written only on the fused
// path; otherwise it is left where the V-cycle put it (no copy).  r32_ready:
// the level-0 FP32 right-hand side already holds fl32(r); first_ready: and
// the level-0 smoother's first step from zero is in its x (and d).
PrecOut precondition_pcg(Handle& h, const double* r, double* z, int k, double* partials, bool r32_ready,
                         bool first_ready) {
  hipStream_t s = h.stream;
  Level& L0 = *h.H.levels[0];
  const int64_t m = h.A.n * k;
  PrecOut o;
  const int type = cycle_type(h);
  if (h.H.precision == 0) {
    o.z64 = vcycle<double>(h, 0, r, k, nullptr, false, type);
    return o;
  }
  if (fused_level0(h, k, partials != nullptr)) {
    if (h.H.precision == 2 && h.H.ah.nslices > 0) {
      o.nb = vcycle0_half(h, r, z, partials);
    } else {
      Level0Io io;
      io.r64 = r;
      io.b32 = L0.b32.get();
      io.z64 = z;
      io.partials = partials;
      float* res = vcycle<float>(h, 0, L0.b32.get(), 1, &io, first_ready, type);
      if (res || io.nb == 0) fail(FEM_ERR_STATE, "fused level-0 smoother did not write its output");
      o.nb = io.nb;
    }
    o.z64 = z;
    return o;
  }
  if (!r32_ready) {
    convert_kernel<double, float><<<grid_for(m, 256), 256, 0, s>>>(m, r, L0.b32.get());
    HIP_LAUNCH_CHECK();
  }
  // k > 1: the last level-0 smoothing step also writes the r.z partials
  // (when it runs on the sliced copy; o.nb stays 0 otherwise)
  Level0Io io;
  io.r64 = r;
  io.partials = partials;
  o.z32 = vcycle<float>(h, 0, L0.b32.get(), k, (partials && k > 1) ? &io : nullptr, r32_ready && first_ready,
                        type);
  o.nb = io.nb;
  return o;
}

// z = M r into z (FP64); returns the count of r.z partials written (fused path)
unsigned precondition(Handle& h, const double* r, double* z, int k, double* partials = nullptr) {
  hipStream_t s = h.stream;
  const int64_t m = h.A.n * k;
  const PrecOut o = precondition_pcg(h, r, z, k, partials, false, false);
  if (o.z32) {
    convert_kernel<float, double><<<grid_for(m, 256), 256, 0, s>>>(m, o.z32, z);
    HIP_LAUNCH_CHECK();
  } else if (o.z64 != z) {
    copy_kernel<<<grid_for(m, 256), 256, 0, s>>>(m, o.z64, z);
    HIP_LAUNCH_CHECK();
  }
  return o.nb;
}

}  // namespace

namespace {
hipEvent_t take_event(Handle& h) {
  if (!h.ev_pool.empty()) {
    hipEvent_t e = h.ev_pool.back();
    h.ev_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  HIP_CHECK(hipEventCreate(&e));
  return e;
}
}  // namespace

ProfScope::ProfScope(Handle& hh, const char* f, double b) : h(hh), fam(f), bytes(b), on(hh.profiling && f && f[0]) {
  if (!on) return;
  a = take_event(h);
  HIP_CHECK(hipEventRecord(a, h.stream));
}

ProfScope::~ProfScope() {
  if (!on) return;
  hipEvent_t b = take_event(h);
  if (hipEventRecord(b, h.stream) != hipSuccess) return;
  h.pending.push_back(PendingTiming{fam, a, b, bytes});
  if (h.pending.size() >= 8192) flush_profile(h);
}

void flush_profile(Handle& h) {
  if (h.pending.empty()) return;
  HIP_CHECK(hipStreamSynchronize(h.stream));
  for (auto& p : h.pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      KernelStat& st = h.stats[p.fam];
      st.launches += 1;
      st.total_ms += ms;
      st.bytes = p.bytes;
    }
    h.ev_pool.push_back(p.a);
    h.ev_pool.push_back(p.b);
  }
  h.pending.clear();
}

void release_profile_events(Handle& h) {
  for (auto& p : h.pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  h.pending.clear();
  for (auto e : h.ev_pool) (void)hipEventDestroy(e);
  h.ev_pool.clear();
}

void spmv64(const Csr& A, const double* x, double* y, hipStream_t s) {
  spmv_op<double, kAssign>(A, 1, x, y, nullptr, (const double*)nullptr, nullptr, nullptr, 0.0, 0.0, s);
}

void apply_operator_int(Handle& h, const double* x, double* y, int k) {
  spmv_op<double, kAssign>(h.A, k, x, y, nullptr, (const double*)nullptr, nullptr, nullptr, 0.0, 0.0, h.stream);
}

const Csr& level_matrix_of(const Handle& h, int lvl, int which) {
  const Hierarchy& H = h.H;
  if (lvl < 0 || lvl >= (int)H.levels.size()) fail(FEM_ERR_ARG, "level out of range");
  if (which == 0 || which == 3) return level_A(h, lvl);
  if (which != 1 && which != 2) fail(FEM_ERR_ARG, "which: 0 operator, 1 prolongator, 2 restriction");
  if (lvl == H.coarse) fail(FEM_ERR_ARG, "the coarsest level has no prolongator or restriction");
  return which == 1 ? H.levels[lvl]->P : H.levels[lvl]->R;
}

void level_smoother_info(const Handle& h, int lvl, double* rho, double* lmin, double* lmax, double* omega) {
  const Hierarchy& H = h.H;
  if (lvl < 0 || lvl >= (int)H.levels.size()) fail(FEM_ERR_ARG, "level out of range");
  if (lvl == H.coarse) fail(FEM_ERR_ARG, "the coarsest level is solved directly, it has no smoother");
  const Level& L = *H.levels[lvl];
  const Smoother S = smoother_for(h, L);
  if (rho) *rho = L.rho;
  if (lmin) *lmin = S.lmin;
  if (lmax) *lmax = S.lmax;
  if (omega) *omega = S.omega;
}

namespace {
template <typename T>
void smoother_step_t(Handle& h, const Csr& A, const T* dinv, int mode, double c1, double c2, const T* x, const T* b,
                     T* d, T* y) {
  if (mode == 0)
    spmv_op<T, kJacobi>(A, 1, x, nullptr, b, dinv, nullptr, y, (T)c1, T(0), h.stream);
  else
    spmv_op<T, kCheb>(A, 1, x, nullptr, b, dinv, d, y, (T)c1, (T)c2, h.stream);
}
}  // namespace

void smoother_step_int(Handle& h, int lvl, int mode, double c1, double c2, const double* x, const double* b,
                       double* d, double* y) {
  const Hierarchy& H = h.H;
  if (lvl < 0 || lvl >= (int)H.levels.size() || lvl == H.coarse) fail(FEM_ERR_ARG, "level has no smoother");
  if (mode != 0 && mode != 1) fail(FEM_ERR_ARG, "mode: 0 Jacobi, 1 Chebyshev");
  const Csr& A = level_A(h, lvl);
  const Level& L = *H.levels[lvl];
  hipStream_t s = h.stream;
  const int64_t n = A.n;
  if (H.precision == 0) {
    smoother_step_t<double>(h, A, L.dinv64.get(), mode, c1, c2, x, b, d, y);
  } else {
    if (!A.v32.size() || (A.nslices > 0 && !A.s32.size()) || !L.dinv32.size())
      fail(FEM_ERR_STATE, "no FP32 copy of this level");
    DevBuf<float> xf(n), bf(n), df(n), yf(n);
    convert_kernel<double, float><<<grid_for(n, 256), 256, 0, s>>>(n, x, xf.get());
    convert_kernel<double, float><<<grid_for(n, 256), 256, 0, s>>>(n, b, bf.get());
    convert_kernel<double, float><<<grid_for(n, 256), 256, 0, s>>>(n, d, df.get());
    HIP_LAUNCH_CHECK();
    smoother_step_t<float>(h, A, L.dinv32.get(), mode, c1, c2, xf.get(), bf.get(), df.get(), yf.get());
    convert_kernel<float, double><<<grid_for(n, 256), 256, 0, s>>>(n, yf.get(), y);
    convert_kernel<float, double><<<grid_for(n, 256), 256, 0, s>>>(n, df.get(), d);
    HIP_LAUNCH_CHECK();
    HIP_CHECK(hipStreamSynchronize(s));
  }
}

void apply_level_int(Handle& h, int lvl, int which, const double* x, double* y) {
  hipStream_t s = h.stream;
  if (which == 3) {
    // the coarsest-level solve the V-cycle applies (FP64 inverse in every
    // precision; b and x in the level's vector type)
    const Hierarchy& H = h.H;
    if (lvl != H.coarse) fail(FEM_ERR_ARG, "which 3 (coarse solve) needs the coarsest level");
    const int64_t nc = H.nc;
    if (H.precision == 0) {
      coarse_kernel<double><<<(unsigned)nc, 64, 0, s>>>(nc, 1, H.cinv64.get(), x, y);
      HIP_LAUNCH_CHECK();
    } else {
      DevBuf<float> bf(nc), xf(nc);
      convert_kernel<double, float><<<grid_for(nc, 256), 256, 0, s>>>(nc, x, bf.get());
      coarse_kernel<float><<<(unsigned)nc, 64, 0, s>>>(nc, 1, H.cinv64.get(), bf.get(), xf.get());
      convert_kernel<float, double><<<grid_for(nc, 256), 256, 0, s>>>(nc, xf.get(), y);
      HIP_LAUNCH_CHECK();
      HIP_CHECK(hipStreamSynchronize(s));
    }
    return;
  }
  const Csr& M = level_matrix_of(h, lvl, which);
  if (h.H.precision == 0) {
    spmv_op<double, kAssign>(M, 1, x, y, nullptr, (const double*)nullptr, nullptr, nullptr, 0.0, 0.0, s);
  } else {
    if (!M.v32.size() || (M.nslices > 0 && !M.s32.size())) fail(FEM_ERR_STATE, "no FP32 copy of this matrix");
    DevBuf<float> xf(M.m > 0 ? M.m : 1), yf(M.n > 0 ? M.n : 1);
    if (M.m > 0) {
      convert_kernel<double, float><<<grid_for(M.m, 256), 256, 0, s>>>(M.m, x, xf.get());
      HIP_LAUNCH_CHECK();
    }
    spmv_op<float, kAssign>(M, 1, xf.get(), yf.get(), nullptr, (const float*)nullptr, nullptr, nullptr, 0.0f, 0.0f, s);
    if (M.n > 0) {
      convert_kernel<float, double><<<grid_for(M.n, 256), 256, 0, s>>>(M.n, yf.get(), y);
      HIP_LAUNCH_CHECK();
    }
    HIP_CHECK(hipStreamSynchronize(s));
  }
}

void ensure_krylov(Handle& h, int k) {
  const int64_t n = h.A.n;
  if (h.kcap < k) {
    ++h.gen;
    h.kx.resize(n * k); h.kr.resize(n * k); h.kp.resize(n * k); h.kq.resize(n * k);
    h.kz.resize(n * k); h.kb.resize(n * k);
    h.ubuf.resize(n * k); h.ubuf2.resize(n * k);
    h.partials.resize((size_t)std::max<int64_t>(4096, n / 256 + 2) * 64);
    h.scal.resize((size_t)S_COUNT * 64);
    if (!h.h_scal) HIP_CHECK(hipHostMalloc((void**)&h.h_scal, sizeof(double) * S_COUNT * 64, hipHostMallocDefault));
    if (!h.h_act) {
      HIP_CHECK(hipHostMalloc((void**)&h.h_act, sizeof(double) * 2 * kActRec,
                              hipHostMallocMapped | hipHostMallocCoherent));
      HIP_CHECK(hipHostGetDevicePointer((void**)&h.d_act, h.h_act, 0));
    }
    h.kcap = k;
  }
  for (auto& Lp : h.H.levels) {
    Level& L = *Lp;
    if (L.wk >= k) continue;
    ++h.gen;
    const int64_t nl = (&L == h.H.levels[0].get()) ? h.A.n : L.A.n;
    if (h.H.precision == 0) {
      L.x64.resize(nl * k); L.t64.resize(nl * k); L.b64.resize(nl * k); L.r64.resize(nl * k); L.d64.resize(nl * k);
    } else {
      L.x32.resize(nl * k); L.t32.resize(nl * k); L.b32.resize(nl * k); L.r32.resize(nl * k); L.d32.resize(nl * k);
    }
    L.wk = k;
  }
}

// the preconditioner as PCG applies it (for one column: the fused level-0 ends
// and, in precision 2, the FP16 level-0 smoother; the r.z partials are unused)
void apply_vcycle_fp64(Handle& h, const double* r, double* z) { precondition(h, r, z, 1, h.partials.get()); }

namespace {

// Runs `body` (a fixed sequence of launches on h.stream that reads and writes
// device memory only) either directly or as a HIP graph captured at its first
// use and replayed while the buffers, hierarchy, k and tolerance it was
// captured with are unchanged.  Replays launch the same kernels with the same
// arguments, so results are bitwise those of the direct launches.
template <typename Fn>
void run_iteration(Handle& h, int k, double tol2, Fn&& body) {
  if (h.prm.launchMode == 0 || h.profiling || h.stream == nullptr) {
    body();
    return;
  }
  if (h.iter_graph && (h.graph_gen != h.gen || h.graph_k != k || h.graph_tol2 != tol2)) {
    HIP_CHECK(hipGraphExecDestroy(h.iter_graph));
    h.iter_graph = nullptr;
  }
  if (!h.iter_graph) {
    HIP_CHECK(hipStreamBeginCapture(h.stream, hipStreamCaptureModeRelaxed));
    try {
      body();
    } catch (...) {
      hipGraph_t junk = nullptr;
      (void)hipStreamEndCapture(h.stream, &junk);
      if (junk) (void)hipGraphDestroy(junk);
      throw;
    }
    hipGraph_t g = nullptr;
    HIP_CHECK(hipStreamEndCapture(h.stream, &g));
    hipError_t e = hipGraphInstantiate(&h.iter_graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIP_CHECK(e);
    h.graph_gen = h.gen;
    h.graph_k = k;
    h.graph_tol2 = tol2;
  }
  HIP_CHECK(hipGraphLaunch(h.iter_graph, h.stream));
}

void finish(Handle& h, int nb, int k, int mode, int slot, double tol2 = 0.0) {
  pcg_finish_kernel<<<1, kFinNT, 0, h.stream>>>(nb, k, h.partials.get(), h.scal.get(), mode, slot, tol2,
                                                 mode == F_RR ? h.d_act : nullptr);
  HIP_LAUNCH_CHECK();
}

template <int KP, typename B>
void launch_dot(Handle& h, int k, const double* a, const B* b, int mode, int slot) {
  const int64_t n = h.A.n;
  unsigned g = vec_grid(n, KP);
  pcg_dot_kernel<KP, B><<<g, kNT, 0, h.stream>>>(n, k, a, b, h.partials.get());
  HIP_LAUNCH_CHECK();
  finish(h, (int)g, k, mode, slot);
}

// mode F_STORE: scal[slot] = a.b; F_BETA: beta = a.b / rz, rz = a.b
template <typename B>
void dot(Handle& h, int k, const double* a, const B* b, int mode, int slot) {
  switch (pow2_at_least(k)) {
    case 1: launch_dot<1, B>(h, k, a, b, mode, slot); break;
    case 2: launch_dot<2, B>(h, k, a, b, mode, slot); break;
    case 4: launch_dot<4, B>(h, k, a, b, mode, slot); break;
    case 8: launch_dot<8, B>(h, k, a, b, mode, slot); break;
    case 16: launch_dot<16, B>(h, k, a, b, mode, slot); break;
    case 32: launch_dot<32, B>(h, k, a, b, mode, slot); break;
    default: launch_dot<64, B>(h, k, a, b, mode, slot); break;
  }
}

template <int KP>
void launch_update(Handle& h, int k, double tol2, float* r32, const First0* f0) {
  const int64_t n = h.A.n;
  unsigned g = vec_grid(n, KP);
  pcg_update_kernel<KP><<<g, kNT, 0, h.stream>>>(n, k, h.kx.get(), h.kr.get(), h.kp.get(), h.kq.get(),
                                                 h.partials.get(), h.scal.get(), f0 ? f0->b32 : r32,
                                                 f0 ? f0->dinv : nullptr, f0 ? f0->x0 : nullptr,
                                                 f0 ? f0->d0 : nullptr, f0 ? f0->c0 : 0.0f);
  HIP_LAUNCH_CHECK();
  finish(h, (int)g, k, F_RR, 0, tol2);
}

// x += alpha p, r -= alpha q, |r|^2, and r32 = fl32(r) when r32 is given
// (with f0: fl32(r) into the level-0 right-hand side and the smoother's first
// step from zero into its x and d)
void update(Handle& h, int k, double tol2, float* r32, const First0* f0 = nullptr) {
  switch (pow2_at_least(k)) {
    case 1: launch_update<1>(h, k, tol2, r32, f0); break;
    case 2: launch_update<2>(h, k, tol2, r32, f0); break;
    case 4: launch_update<4>(h, k, tol2, r32, f0); break;
    case 8: launch_update<8>(h, k, tol2, r32, f0); break;
    case 16: launch_update<16>(h, k, tol2, r32, f0); break;
    case 32: launch_update<32>(h, k, tol2, r32, f0); break;
    default: launch_update<64>(h, k, tol2, r32, f0); break;
  }
}

void launch_sell_dot(Handle& h) {
  const Csr& A = h.A;
  if (A.srow.size() || A.sell_g != 1)
    fail(FEM_ERR_STATE, "the operator's sliced-ELL copy must keep the row order with one lane per row");
  if (A.lds_bs > 0) {
    const unsigned nb = (unsigned)((A.nslices + A.lds_bs - 1) / A.lds_bs);
    pcg_sell_lds_dot_kernel<<<nb, kNT, sizeof(double) * (size_t)A.lds_maxc, h.stream>>>(
        A.n, A.nslices, sell_view(A), lds_view(A), A.s64.get(), h.kp.get(), h.kq.get(), h.partials.get());
    HIP_LAUNCH_CHECK();
    finish(h, (int)nb, 1, F_ALPHA, 0);
    return;
  }
  unsigned g = grid_for(A.nslices, kNT / 64);
  pcg_sell_dot_kernel<<<g, kNT, 0, h.stream>>>(A.n, A.nslices, sell_view(A), A.s64.get(), h.kp.get(), h.kq.get(),
                                               h.partials.get());
  HIP_LAUNCH_CHECK();
  finish(h, (int)g, 1, F_ALPHA, 0);
}

template <int G>
void launch_spmv_dot(Handle& h) {
  const Csr& A = h.A;
  int64_t g = (A.n * G + kNT - 1) / kNT;
  if (g > 4096) g = 4096;
  pcg_spmv_dot_kernel<G><<<(unsigned)g, kNT, 0, h.stream>>>(A.n, A.rp.get(), A.ci.get(), A.v64.get(), h.kp.get(),
                                                            h.kq.get(), h.partials.get());
  HIP_LAUNCH_CHECK();
  finish(h, (int)g, 1, F_ALPHA, 0);
}

// FP64 columns per lane of the multi-column sliced operator + dot (as sell_multi_cols)
constexpr int kDotCols = 4;

template <int KP>
void launch_spmm_dot(Handle& h, int k) {
  const Csr& A = h.A;
  int64_t g = (A.n * KP + kNT - 1) / kNT;
  if (g > 4096) g = 4096;
  if (k % kDotCols == 0 && A.nslices > 0 && A.sell_g == 1 && A.s64.size() && A.srow.size() == 0) {
    constexpr int L = KP >= kDotCols ? KP / kDotCols : 1;
    int64_t gs = (A.nslices * L + kNT / 64 - 1) / (kNT / 64);
    if (gs > 4096) gs = 4096;
    g = gs;
    pcg_sell_multi_dot_kernel<L, kDotCols><<<(unsigned)g, kNT, 0, h.stream>>>(
        A.n, k, A.nslices, sell_view(A), A.s64.get(), h.kp.get(), h.kq.get(), h.partials.get());
  } else if (k % 2 == 0) {
    // KP >= k, so KP / 2 lanes per row cover the k / 2 column pairs
    pcg_spmm_dot_vec_kernel<(KP >= 2 ? KP / 2 : 1)><<<(unsigned)g, kNT, 0, h.stream>>>(
        A.n, k, A.rp.get(), A.ci.get(), A.v64.get(), h.kp.get(), h.kq.get(), h.partials.get());
  } else {
    pcg_spmm_dot_kernel<KP><<<(unsigned)g, kNT, 0, h.stream>>>(A.n, k, A.rp.get(), A.ci.get(), A.v64.get(),
                                                               h.kp.get(), h.kq.get(), h.partials.get());
  }
  HIP_LAUNCH_CHECK();
  finish(h, (int)g, k, F_ALPHA, 0);
}

void spmv_dot(Handle& h, int k) {
  const Csr& A = h.A;
  ProfScope ps(h, "pcg_spmv", matrix_bytes<double>(A, k) + 3.0 * 8 * A.n * k);
  if (k == 1 && A.nslices > 0 && A.s64.size() && A.sell_g == 1 && A.srow.size() == 0) {
    launch_sell_dot(h);
  } else if (k == 1) {
    switch (lanes_for(A)) {
      case 4: launch_spmv_dot<4>(h); break;
      case 8: launch_spmv_dot<8>(h); break;
      case 16: launch_spmv_dot<16>(h); break;
      case 32: launch_spmv_dot<32>(h); break;
      default: launch_spmv_dot<64>(h); break;
    }
  } else {
    switch (pow2_at_least(k)) {
      case 2: launch_spmm_dot<2>(h, k); break;
      case 4: launch_spmm_dot<4>(h, k); break;
      case 8: launch_spmm_dot<8>(h, k); break;
      case 16: launch_spmm_dot<16>(h, k); break;
      case 32: launch_spmm_dot<32>(h, k); break;
      default: launch_spmm_dot<64>(h, k); break;
    }
  }
}

void set_scal(Handle& h, int slot, int k, const double* v) {
  HIP_CHECK(hipMemcpyAsync(h.scal.get() + slot * k, v, sizeof(double) * k, hipMemcpyHostToDevice, h.stream));
}

// Convergence record `seq` of the |r|^2 finish (act flags and |r|^2 per
// column), read from host-mapped memory once event `ev` (recorded after that
// finish) has completed; polled, so there is no sleep / wake-up in between.
void read_record(Handle& h, hipEvent_t ev, uint64_t seq, int k, double* act, double* rr) {
  for (;;) {
    const hipError_t e = hipEventQuery(ev);
    if (e == hipSuccess) break;
    if (e != hipErrorNotReady) HIP_CHECK(e);
  }
  const volatile double* rec = h.h_act + (seq & 1) * kActRec;
  if (rec[128] != (double)seq) fail(FEM_ERR_STATE, "convergence record out of sequence");
  for (int c = 0; c < k; ++c) {
    act[c] = rec[c];
    rr[c] = rec[64 + c];
  }
}

void get_scal(Handle& h, int slot, int k, double* v) {
  HIP_CHECK(hipMemcpyAsync(h.h_scal, h.scal.get() + slot * k, sizeof(double) * k, hipMemcpyDeviceToHost, h.stream));
  HIP_CHECK(hipStreamSynchronize(h.stream));
  std::memcpy(v, h.h_scal, sizeof(double) * k);
}

}  // namespace

// Solve A x = b for k columns held in h.kb / h.kx (internal numbering).
int pcg_solve(Handle& h, int k, fem_solve_info* info) {
  hipStream_t s = h.stream;
  const int64_t n = h.A.n;
  const int64_t m = n * k;
  const double tol = h.prm.tolerance;
  const double tol2 = tol * tol;
  std::vector<double> bb(k), rr(k), act(k), one(k, 1.0);

  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  HIP_CHECK(hipEventRecord(e0, s));

  dot(h, k, h.kb.get(), h.kb.get(), F_STORE, S_BB);
  get_scal(h, S_BB, k, bb.data());

  int iters = 0;
  int status = FEM_OK;
  auto true_residual = [&](std::vector<double>& rel) {
    // kq = b - A x ; rel_c = ||kq_c|| / ||b_c||
    spmv_op<double, kResid>(h.A, k, h.kx.get(), h.kq.get(), h.kb.get(), (const double*)nullptr, nullptr, nullptr,
                            0.0, 0.0, s);
    dot(h, k, h.kq.get(), h.kq.get(), F_STORE, S_TRUE);
    std::vector<double> t(k);
    get_scal(h, S_TRUE, k, t.data());
    rel.resize(k);
    for (int c = 0; c < k; ++c) rel[c] = bb[c] > 0 ? std::sqrt(t[c] / bb[c]) : 0.0;
  };

  std::vector<double> rel(k, 0.0);
  if (h.prm.solverType == 0) {
    // one V-cycle: x = M b (the initial guess is not used, as in the reference)
    precondition(h, h.kb.get(), h.kx.get(), k);
    iters = 1;
    true_residual(rel);
    for (int c = 0; c < k; ++c)
      if (rel[c] > tol) status = FEM_ERR_NOT_CONVERGED;
  } else {
    // a zero right-hand side has the exact solution x = 0, whatever the
    // initial guess (its column then starts, and stays, converged)
    unsigned long long zmask = 0;
    for (int c = 0; c < k; ++c)
      if (!(bb[c] > 0)) zmask |= 1ull << c;
    if (zmask) {
      zero_cols_kernel<<<grid_for(m, 256), 256, 0, s>>>(m, k, zmask, h.kx.get());
      HIP_LAUNCH_CHECK();
    }
    // r = b - A x0
    spmv_op<double, kResid>(h.A, k, h.kx.get(), h.kr.get(), h.kb.get(), (const double*)nullptr, nullptr, nullptr,
                            0.0, 0.0, s);
    dot(h, k, h.kr.get(), h.kr.get(), F_STORE, S_RR);
    get_scal(h, S_RR, k, rr.data());
    bool any = false;
    for (int c = 0; c < k; ++c) {
      act[c] = (bb[c] > 0 && rr[c] > tol2 * bb[c]) ? 1.0 : 0.0;
      any = any || act[c] != 0.0;
    }
    set_scal(h, S_ACT, k, act.data());
    const std::vector<double> zero_seq(k, 0.0);
    set_scal(h, S_SEQ, k, zero_seq.data());  // |r|^2 finishes of this solve are numbered from 1
    uint64_t seq_issued = 0;
    bool ahead = false;
    hipEvent_t evs[2];
    HIP_CHECK(hipEventCreateWithFlags(&evs[0], hipEventDisableTiming));
    HIP_CHECK(hipEventCreateWithFlags(&evs[1], hipEventDisableTiming));
    struct EvGuard {
      hipEvent_t* e;
      ~EvGuard() {
        (void)hipEventDestroy(e[0]);
        (void)hipEventDestroy(e[1]);
      }
    } ev_guard{evs};
    int restarts = 0;
    // the unfused mixed-precision V-cycle takes fl32(r), written by the update
    float* r32 = (h.H.precision >= 1 && !fused_level0(h, k, true)) ? h.H.levels[0]->b32.get() : nullptr;
    // and the level-0 smoother's first step from zero (with fl32(r) on the
    // fused path) comes from the update as well
    const First0 f0 = first0_for(h, k);
    while (any) {
      // z = M r and rz = r.z (fused into the last level-0 smoothing step when
      // possible); z stays where the preconditioner left it
      auto precondition_dot = [&](int mode, int slot, bool r32_ready) {
        const PrecOut o = precondition_pcg(h, h.kr.get(), h.kz.get(), k, h.partials.get(),
                                           r32_ready && (r32 || f0.on), r32_ready && f0.on);
        if (o.nb) finish(h, (int)o.nb, k, mode, slot);
        else if (o.z32) dot(h, k, h.kr.get(), o.z32, mode, slot);
        else dot(h, k, h.kr.get(), o.z64, mode, slot);
        return o;
      };
      const PrecOut o0 = precondition_dot(F_STORE, S_RZ, false);
      if (o0.z32) convert_kernel<float, double><<<grid_for(m, 256), 256, 0, s>>>(m, o0.z32, h.kp.get());
      else copy_kernel<<<grid_for(m, 256), 256, 0, s>>>(m, o0.z64, h.kp.get());
      HIP_LAUNCH_CHECK();
      // the first iteration starts from p = z; every later one is the fixed
      // sequence below (captured once, then replayed), which follows an update
      // (so fl32(r) is in place)
      // an F-cycle is not a symmetric preconditioner: its beta is the flexible one
      const bool flexible = cycle_type(h) == kCycleF;
      auto iteration = [&]() {
        // z, beta, rz
        const PrecOut o = flexible ? precondition_dot(F_STORE, S_RZN, true) : precondition_dot(F_BETA, 0, true);
        if (flexible) {
          if (o.z32) dot(h, k, h.kq.get(), o.z32, F_FLEX, 0);
          else dot(h, k, h.kq.get(), o.z64, F_FLEX, 0);
        }
        if (o.z32)
          pcg_p_kernel<float><<<grid_for(m, 256), 256, 0, s>>>(m, k, o.z32, h.kp.get(), h.scal.get());
        else
          pcg_p_kernel<double><<<grid_for(m, 256), 256, 0, s>>>(m, k, o.z64, h.kp.get(), h.scal.get());
        HIP_LAUNCH_CHECK();
        spmv_dot(h, k);           // q = A p, alpha
        update(h, k, tol2, r32, f0.on ? &f0 : nullptr);  // x, r, rr, act
      };
      // Iteration n ends with the n-th |r|^2 finish of this solve; the host
      // reads its record while iteration n + 1 may already run: the next
      // iteration is queued ahead whenever the residual history predicts
      // more than two further iterations.  An iteration queued after the
      // last needed one leaves x and r unchanged (alpha = beta = 0 for
      // converged columns).
      const uint64_t seq0 = seq_issued;
      auto issue = [&](bool first) {
        if (first) {
          spmv_dot(h, k);
          update(h, k, tol2, r32, f0.on ? &f0 : nullptr);
        } else {
          run_iteration(h, k, tol2, iteration);
        }
        ++seq_issued;
        HIP_CHECK(hipEventRecord(evs[seq_issued & 1], s));
      };
      if (iters < h.prm.maxIters) issue(true);
      uint64_t seq_done = seq0;
      std::vector<double> rr_prev(k, -1.0), rr_prev2(k, -1.0), rrn(k);
      while (seq_done < seq_issued) {
        const bool room = iters + (int)(seq_issued - seq_done) < h.prm.maxIters;
        if (ahead && room && seq_issued == seq_done + 1) issue(false);
        read_record(h, evs[(seq_done + 1) & 1], seq_done + 1, k, act.data(), rrn.data());
        ++seq_done;
        ++iters;
        bool still = false;
        double remaining = 0.0;
        for (int c = 0; c < k; ++c) {
          if (act[c] == 0.0) continue;
          still = true;
          // iterations left at the geometric rate of the last two
          double rem = 1e30;
          if (rr_prev2[c] > 0.0 && rrn[c] > 0.0 && rrn[c] < rr_prev2[c])
            rem = std::log(tol2 * bb[c] / rrn[c]) / (0.5 * std::log(rrn[c] / rr_prev2[c]));
          remaining = std::max(remaining, rem);
          rr_prev2[c] = rr_prev[c];
          rr_prev[c] = rrn[c];
        }
        if (!still || iters >= h.prm.maxIters) break;
        ahead = remaining > 2.0;
        if (seq_issued == seq_done) issue(false);
      }
      // verify on the true residual; restart columns that drifted
      true_residual(rel);
      any = false;
      for (int c = 0; c < k; ++c) {
        act[c] = (rel[c] > tol) ? 1.0 : 0.0;
        any = any || act[c] != 0.0;
      }
      if (!any || iters >= h.prm.maxIters || restarts >= 3) break;
      ++restarts;
      copy_kernel<<<grid_for(m, 256), 256, 0, s>>>(m, h.kq.get(), h.kr.get());
      HIP_LAUNCH_CHECK();
      set_scal(h, S_ACT, k, act.data());
    }
    if (any) status = FEM_ERR_NOT_CONVERGED;
    if (rel.empty() || iters == 0) true_residual(rel);
    for (int c = 0; c < k; ++c)
      if (rel[c] > tol) status = FEM_ERR_NOT_CONVERGED;
  }
  HIP_CHECK(hipEventRecord(e1, s));
  HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (info) {
    info->iterations = iters;
    double mx = 0.0;
    for (int c = 0; c < k; ++c) mx = std::max(mx, rel[c]);
    info->rel_residual = mx;
    info->converged = (status == FEM_OK) ? 1 : 0;
    info->solve_ms = ms;
  }
  return status;
}

// ---- allocation-free preconditioner application ---------------------------------
// Device scratch of apply_preconditioner_noalloc: r and z in internal
// numbering (kNoAllocRows doubles each), then the r.z partials of the fused
// level-0 smoother's last step (at most one per 64-row slice).  The array is
// one copy per GPU, held until the process ends, and shared by every call on
// that GPU.  Its name is a host-side stand-in in host code: kernels receive
// the per-device address from hipGetSymbolAddress.
constexpr int64_t kNoAllocRows = (int64_t)1 << 21;
constexpr int64_t kNoAllocPartials = kNoAllocRows / 64 + 64;
static __device__ __align__(256) double scratch[2 * kNoAllocRows + kNoAllocPartials];

namespace {
std::mutex g_scratch_mu;
std::map<int, double*> g_scratch_addr;  // device ordinal -> address of scratch

// the address of scratch on device dev, which must be the current device
// (looked up once per device)
double* scratch_address(int dev) {
  int cur = -1;
  HIP_CHECK(hipGetDevice(&cur));
  if (cur != dev)
    fail(FEM_ERR_STATE, "the current HIP device (" + std::to_string(cur) + ") is not the handle's device (" +
                            std::to_string(dev) + ")");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  auto it = g_scratch_addr.find(dev);
  if (it != g_scratch_addr.end()) return it->second;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || p == nullptr)
    fail(FEM_ERR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
  g_scratch_addr[dev] = static_cast<double*>(p);
  return static_cast<double*>(p);
}
}  // namespace

int64_t noalloc_max_rows() { return kNoAllocRows; }

void noalloc_init(const Handle& h) {
  int cur = -1;
  if (hipGetDevice(&cur) == hipSuccess && cur == h.device) (void)scratch_address(h.device);
}

void apply_preconditioner_noalloc(Handle& h, const double* r_user, double* z_user) {
  const int64_t n = h.A.n;
  if (n == 0) fail(FEM_ERR_EMPTY, "Error no matrix specified");
  if (n > kNoAllocRows)
    fail(FEM_ERR_CAPACITY, "fem_apply_preconditioner_noalloc: " + std::to_string(n) +
                               " rows exceed its limit of " + std::to_string(kNoAllocRows) + " rows");
  if (!h.H.built) fail(FEM_ERR_STATE, "fem_apply_preconditioner_noalloc: no hierarchy; call fem_setup first");
  bool ready = h.kcap >= 1;
  for (auto& Lp : h.H.levels) ready = ready && Lp->wk >= 1;
  if (!ready) fail(FEM_ERR_STATE, "fem_apply_preconditioner_noalloc: work vectors missing; call fem_setup first");
  if (h.profiling) fail(FEM_ERR_STATE, "fem_apply_preconditioner_noalloc: profiling is enabled");
  double* base = scratch_address(h.device);
  double* r = base;
  double* z = base + kNoAllocRows;
  double* partials = base + 2 * kNoAllocRows;
  gather_user_to_int(h, r_user, r, 1);
  precondition(h, r, z, 1, partials);
  scatter_int_to_user(h, z, z_user, 1);
}

}  // namespace fem
