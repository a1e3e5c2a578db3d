// Device sort / scan / fill primitives for the setup phase (rocPRIM).
// Radix sorts are stable, and integer scans are exact, so everything built on
// them is deterministic.
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "fem_internal.hpp"

namespace fem {

namespace {

template <typename K>
void sort_pairs_impl(const K* kin, K* kout, const int32_t* vin, int32_t* vout, size_t n, int end_bit,
                     hipStream_t s) {
  if (n == 0) return;
  size_t bytes = 0;
  HIP_CHECK(rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, n, 0u, (unsigned)end_bit, s));
  DevBuf<char> tmp(bytes > 0 ? bytes : 1);
  HIP_CHECK(rocprim::radix_sort_pairs((void*)tmp.get(), bytes, kin, kout, vin, vout, n, 0u,
                                      (unsigned)end_bit, s));
  HIP_CHECK(hipStreamSynchronize(s));  // tmp is freed on return
}

template <typename T>
__global__ void tail_total_kernel(const T* in, T* out, size_t n) {
  if (threadIdx.x == 0 && blockIdx.x == 0) out[n] = (n ? out[n - 1] + in[n - 1] : T(0));
}

template <typename T>
int64_t exclusive_scan_impl(const T* in, T* out, size_t n, hipStream_t s) {
  if (n > 0) {
    size_t bytes = 0;
    HIP_CHECK(rocprim::exclusive_scan(nullptr, bytes, in, out, T(0), n, rocprim::plus<T>(), s));
    DevBuf<char> tmp(bytes > 0 ? bytes : 1);
    HIP_CHECK(rocprim::exclusive_scan((void*)tmp.get(), bytes, in, out, T(0), n, rocprim::plus<T>(), s));
  }
  tail_total_kernel<T><<<1, 64, 0, s>>>(in, out, n);
  HIP_LAUNCH_CHECK();
  T total = 0;
  HIP_CHECK(hipMemcpyAsync(&total, out + n, sizeof(T), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  return (int64_t)total;
}

template <typename T>
__global__ void fill_kernel(T* p, T v, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

__global__ void iota_kernel(int32_t* p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = (int32_t)i;
}

__global__ void f64_to_f32_kernel(const double* in, float* out, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) out[i] = (float)in[i];
}

unsigned fill_grid(size_t n) {
  size_t g = (n + 255) / 256;
  if (g > 8192) g = 8192;
  if (g < 1) g = 1;
  return (unsigned)g;
}

}  // namespace

void dev_sort_pairs_u64_i32(const uint64_t* kin, uint64_t* kout, const int32_t* vin, int32_t* vout,
                            size_t n, int end_bit, hipStream_t s) {
  sort_pairs_impl<uint64_t>(kin, kout, vin, vout, n, end_bit, s);
}
void dev_sort_pairs_u32_i32(const uint32_t* kin, uint32_t* kout, const int32_t* vin, int32_t* vout,
                            size_t n, int end_bit, hipStream_t s) {
  sort_pairs_impl<uint32_t>(kin, kout, vin, vout, n, end_bit, s);
}
int64_t dev_exclusive_scan_i32(const int32_t* in, int32_t* out, size_t n, hipStream_t s) {
  return exclusive_scan_impl<int32_t>(in, out, n, s);
}
int64_t dev_exclusive_scan_i64(const int64_t* in, int64_t* out, size_t n, hipStream_t s) {
  return exclusive_scan_impl<int64_t>(in, out, n, s);
}
void dev_fill_i32(int32_t* p, int32_t v, size_t n, hipStream_t s) {
  if (!n) return;
  fill_kernel<int32_t><<<fill_grid(n), 256, 0, s>>>(p, v, n);
  HIP_LAUNCH_CHECK();
}
void dev_fill_f64(double* p, double v, size_t n, hipStream_t s) {
  if (!n) return;
  fill_kernel<double><<<fill_grid(n), 256, 0, s>>>(p, v, n);
  HIP_LAUNCH_CHECK();
}
void dev_fill_f32(float* p, float v, size_t n, hipStream_t s) {
  if (!n) return;
  fill_kernel<float><<<fill_grid(n), 256, 0, s>>>(p, v, n);
  HIP_LAUNCH_CHECK();
}
void dev_iota_i32(int32_t* p, size_t n, hipStream_t s) {
  if (!n) return;
  iota_kernel<<<fill_grid(n), 256, 0, s>>>(p, n);
  HIP_LAUNCH_CHECK();
}
void dev_f64_to_f32(const double* in, float* out, size_t n, hipStream_t s) {
  if (!n) return;
  f64_to_f32_kernel<<<fill_grid(n), 256, 0, s>>>(in, out, n);
  HIP_LAUNCH_CHECK();
}

}  // namespace fem
