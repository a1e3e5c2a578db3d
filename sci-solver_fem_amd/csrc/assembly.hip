// Mesh reordering, P1 sparsity pattern and deterministic P1 assembly.
//
// Reference behaviour being replaced:
//   pattern  : TetMesh::need_neighbors + tetmesh2ell / trimesh2ell
//              (src/core/cuda/tetmesh.cu:112-170, src/core/cuda/cutil.cu:169-230,298-348):
//              row i holds i and every vertex sharing an element with i.
//   values   : element_loop_3d_kernel / element_loop_kernel
//              (src/core/cuda/perform_element_loop_3D.cuh:481-627,
//               src/core/cuda/perform_element_loop_2D.cuh:288-372):
//              K_e = co * |T| grad(phi_a).grad(phi_b) + M_e, scattered with FP64 CAS atomics.
//
// Here the scatter is replaced by two passes: one thread per element computes
// the element's entries once (coalesced int4 connectivity and coefficient
// reads), then one thread per matrix row gathers the rows of its incident
// element matrices in ascending element order.  No float atomics, a fixed
// summation order (bit-reproducible), and A is bitwise symmetric because both
// entries of a pair add the same stored element values in the same order.
#include "fem_internal.hpp"
#include "kernels.hpp"

#include <algorithm>
#include <cmath>

namespace fem {

namespace {

constexpr int kRowBlock = 64;   // threads (rows) per block for row-owner kernels
constexpr int kRowCap = 64;     // LDS row capacity; longer rows use the global path
constexpr int kIncChunk = 8;    // incidences whose loads a row-owner thread issues together

__device__ __forceinline__ uint64_t spread3(uint64_t v) {  // 21 bits -> every 3rd bit
  v &= 0x1fffffull;
  v = (v | (v << 32)) & 0x1f00000000ffffull;
  v = (v | (v << 16)) & 0x1f0000ff0000ffull;
  v = (v | (v << 8)) & 0x100f00f00f00f00full;
  v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}
__device__ __forceinline__ uint64_t spread2(uint64_t v) {  // 32 bits -> every 2nd bit
  v &= 0xffffffffull;
  v = (v | (v << 16)) & 0x0000ffff0000ffffull;
  v = (v | (v << 8)) & 0x00ff00ff00ff00ffull;
  v = (v | (v << 4)) & 0x0f0f0f0f0f0f0f0full;
  v = (v | (v << 2)) & 0x3333333333333333ull;
  v = (v | (v << 1)) & 0x5555555555555555ull;
  return v;
}

__global__ void morton_kernel(int64_t nv, int dim, const double* __restrict__ xyz, double x0, double y0,
                              double z0, double sx, double sy, double sz, uint64_t* __restrict__ key,
                              int32_t* __restrict__ idx) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  double x = (xyz[3 * v + 0] - x0) * sx;
  double y = (xyz[3 * v + 1] - y0) * sy;
  double z = (xyz[3 * v + 2] - z0) * sz;
  uint64_t k;
  if (dim == 3) {
    double m = 2097151.0;  // 2^21 - 1
    uint64_t qx = (uint64_t)fmin(fmax(x * m, 0.0), m);
    uint64_t qy = (uint64_t)fmin(fmax(y * m, 0.0), m);
    uint64_t qz = (uint64_t)fmin(fmax(z * m, 0.0), m);
    k = spread3(qx) | (spread3(qy) << 1) | (spread3(qz) << 2);
  } else {
    double m = 4294967295.0;
    uint64_t qx = (uint64_t)fmin(fmax(x * m, 0.0), m);
    uint64_t qy = (uint64_t)fmin(fmax(y * m, 0.0), m);
    k = spread2(qx) | (spread2(qy) << 1);
  }
  key[v] = k;
  idx[v] = (int32_t)v;
}

__global__ void invert_perm_kernel(int64_t n, const int32_t* __restrict__ perm, int32_t* __restrict__ iperm) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) iperm[perm[i]] = (int32_t)i;
}

__global__ void coords_kernel(int64_t nv, const double* __restrict__ xyz, const int32_t* __restrict__ perm,
                              double4* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  int64_t u = perm[i];
  out[i] = make_double4(xyz[3 * u], xyz[3 * u + 1], xyz[3 * u + 2], 0.0);
}

// relabel element vertices and compute the element sort key (lowest vertex)
__global__ void relabel_elems_kernel(int64_t ne, int npe, const int32_t* __restrict__ conn_user,
                                     const int32_t* __restrict__ iperm, int4* __restrict__ conn_tmp,
                                     uint32_t* __restrict__ key, int32_t* __restrict__ idx) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  int v[4] = {-1, -1, -1, -1};
  int mn = 0x7fffffff;
  for (int a = 0; a < npe; ++a) {
    v[a] = iperm[conn_user[npe * e + a]];
    mn = min(mn, v[a]);
  }
  conn_tmp[e] = make_int4(v[0], v[1], v[2], v[3]);
  key[e] = (uint32_t)mn;
  idx[e] = (int32_t)e;
}

__global__ void permute_elems_kernel(int64_t ne, const int32_t* __restrict__ eperm, const int4* __restrict__ conn_tmp,
                                     const double* __restrict__ kappa_user, int4* __restrict__ conn,
                                     double* __restrict__ kappa) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  int64_t o = eperm[e];
  conn[e] = conn_tmp[o];
  kappa[e] = kappa_user ? kappa_user[o] : 1.0;
}

__global__ void v2e_keys_kernel(int64_t ne, int npe, const int4* __restrict__ conn, uint32_t* __restrict__ key,
                                int32_t* __restrict__ val) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  int4 c = conn[e];
  int v[4] = {c.x, c.y, c.z, c.w};
  for (int a = 0; a < npe; ++a) {
    key[npe * e + a] = (uint32_t)v[a];
    val[npe * e + a] = (int32_t)(4 * e + a);
  }
}

// ptr[v] = first position of key v in the sorted key array (lower bound)
__global__ void lower_bound_kernel(int64_t nv, const uint32_t* __restrict__ keys, int64_t nk,
                                   int32_t* __restrict__ ptr) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v > nv) return;
  int64_t lo = 0, hi = nk;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if ((int64_t)keys[mid] < v) lo = mid + 1; else hi = mid;
  }
  ptr[v] = (int32_t)lo;
}

// Insert c into the ascending list buf[0..len); returns the new length (cap ends with -1).
template <typename Buf>
__device__ __forceinline__ int sorted_insert(Buf buf, int len, int cap, int c) {
  int lo = 0, hi = len;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (buf[mid] < c) lo = mid + 1; else hi = mid;
  }
  if (lo < len && buf[lo] == c) return len;
  if (len >= cap) return -1;
  for (int t = len; t > lo; --t) buf[t] = buf[t - 1];
  buf[lo] = c;
  return len + 1;
}

struct LdsRow {
  int* p;
  __device__ __forceinline__ int& operator[](int t) const { return p[t * kRowBlock]; }
};

// Row pattern: sorted unique vertices of all elements incident to row vertex.
// Pass 0 counts (cnt[row], or -1 when the row needs more than kRowCap slots),
// pass 1 writes columns at rp[row].  Rows flagged -1 go through the global path.
// With inc given (no row overflowed, slices laid out by sptr), pass 1 also
// writes the row's assembly incidences (element code, positions of the
// element's vertices in the row, one byte each) from the sorted row in LDS,
// the same values as inc_fill_kernel, which searches the written columns.
template <bool kFill>
__global__ void __launch_bounds__(kRowBlock) pattern_rows_kernel(
    int64_t nv, int npe, const int32_t* __restrict__ v2e_ptr, const int32_t* __restrict__ v2e,
    const int4* __restrict__ conn, int32_t* __restrict__ cnt, const int32_t* __restrict__ rp,
    int32_t* __restrict__ ci, int32_t* __restrict__ stash, const int32_t* __restrict__ sptr = nullptr,
    int2* __restrict__ inc = nullptr) {
  __shared__ int lds[kRowBlock * kRowCap];
  // XCD-contiguous rows: neighbouring rows gather the same elements through one L2
  int64_t row = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (row >= nv) return;
  // rows that overflowed kRowCap in the count pass (marked -1 there; the
  // global-memory count then stores their length) are filled by the global path
  if (kFill && (cnt[row] < 0 || cnt[row] > kRowCap)) return;
  // lane-interleaved rows (entry t of lane j at t * kRowBlock + j): the
  // lanes' searches and shifts address distinct banks
  const LdsRow buf{lds + threadIdx.x};
  // the count pass keeps each sorted row in a stash laid out like the LDS
  // rows (entry t of row r at ((r / 64) * kRowCap + t) * 64 + r % 64, a
  // wavefront's entry t one contiguous run), and the fill pass loads it
  // instead of sorting again
  int32_t* const srow = stash + (row >> 6) * (int64_t)(kRowCap * kRowBlock) + (row & 63);
  const int pb = v2e_ptr[row], pe = v2e_ptr[row + 1];
  int len = 0;
  if (kFill) {
    len = cnt[row];
    for (int t = 0; t < len; ++t) buf[t] = srow[t * kRowBlock];
  } else {
    len = sorted_insert(buf, len, kRowCap, (int)row);
    // the element codes and vertices of kIncChunk incidences are loaded
    // together (independent loads), then inserted in incidence order
    for (int p0 = pb; p0 < pe && len >= 0; p0 += kIncChunk) {
      int4 cc[kIncChunk];
#pragma unroll
      for (int u = 0; u < kIncChunk; ++u) cc[u] = p0 + u < pe ? conn[v2e[p0 + u] >> 2] : make_int4(0, 0, 0, 0);
      for (int u = 0; u < kIncChunk && p0 + u < pe && len >= 0; ++u) {
        const int v[4] = {cc[u].x, cc[u].y, cc[u].z, cc[u].w};
        for (int a = 0; a < npe && len >= 0; ++a) len = sorted_insert(buf, len, kRowCap, v[a]);
      }
    }
  }
  if (!kFill) {
    cnt[row] = len;  // -1 marks overflow
    for (int t = 0; t < len; ++t) srow[t * kRowBlock] = buf[t];
  } else {
    int32_t o = rp[row];
    for (int t = 0; t < len; ++t) ci[o + t] = buf[t];
    if (inc && len >= 0) {
      const int64_t sl = row >> 6;
      const int lane = (int)(row & 63);
      const int k0 = sptr[sl], kn = sptr[sl + 1] - k0;
      // kn == 0: the slice sums in place and has no incidences
      for (int p0 = pb; kn > 0 && p0 < pe; p0 += kIncChunk) {
        int code[kIncChunk];
        int4 cc[kIncChunk];
#pragma unroll
        for (int u = 0; u < kIncChunk; ++u) code[u] = p0 + u < pe ? v2e[p0 + u] : 0;
#pragma unroll
        for (int u = 0; u < kIncChunk; ++u) cc[u] = p0 + u < pe ? conn[code[u] >> 2] : make_int4(0, 0, 0, 0);
        for (int u = 0; u < kIncChunk && p0 + u < pe; ++u) {
          const int v[4] = {cc[u].x, cc[u].y, cc[u].z, cc[u].w};
          unsigned slots = 0;
          for (int b = 0; b < npe; ++b) {
            int lo = 0, hi = len;
            while (lo < hi) {
              const int mid = (lo + hi) >> 1;
              if (buf[mid] < v[b]) lo = mid + 1; else hi = mid;
            }
            slots |= (unsigned)lo << (8 * b);
          }
          inc[((int64_t)(k0 + (p0 + u - pb)) << 6) + lane] = make_int2(code[u], (int)slots);
        }
      }
    }
  }
}

// Global path for long rows: upper bound = 1 + npe * incident elements.
template <bool kFill>
__global__ void pattern_rows_global_kernel(int64_t nlong, const int32_t* __restrict__ long_rows,
                                           const int64_t* __restrict__ scratch_off, int npe,
                                           const int32_t* __restrict__ v2e_ptr,
                                           const int32_t* __restrict__ v2e, const int4* __restrict__ conn,
                                           int* __restrict__ scratch, int32_t* __restrict__ cnt,
                                           const int32_t* __restrict__ rp, int32_t* __restrict__ ci) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nlong) return;
  int row = long_rows[t];
  int* buf = scratch + scratch_off[t];
  int cap = (int)(scratch_off[t + 1] - scratch_off[t]);
  int len = sorted_insert(buf, 0, cap, row);
  for (int p = v2e_ptr[row]; p < v2e_ptr[row + 1]; ++p) {
    int4 c = conn[v2e[p] >> 2];
    int v[4] = {c.x, c.y, c.z, c.w};
    for (int a = 0; a < npe; ++a) len = sorted_insert(buf, len, cap, v[a]);
  }
  if (!kFill) {
    cnt[row] = len;
  } else {
    int32_t o = rp[row];
    for (int q = 0; q < len; ++q) ci[o + q] = buf[q];
  }
}

__global__ void collect_long_rows_kernel(int64_t nv, const int32_t* __restrict__ cnt, const int32_t* __restrict__ v2e_ptr,
                                         int npe, int32_t* __restrict__ long_rows, int64_t* __restrict__ cap,
                                         unsigned* __restrict__ nlong) {
  int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= nv) return;
  if (cnt[row] < 0) {
    unsigned slot = atomicAdd(nlong, 1u);
    long_rows[slot] = (int32_t)row;
    cap[slot] = 1 + (int64_t)npe * (v2e_ptr[row + 1] - v2e_ptr[row]);
  }
}

// ---- element geometry --------------------------------------------------------
// P1 gradients of the barycentric basis and the element measure.
// 3D: e_k = p_k - p_0, det = e1 . (e2 x e3) = 6 vol (signed),
//     grad phi_1 = (e2 x e3)/det, grad phi_2 = (e3 x e1)/det, grad phi_3 = (e1 x e2)/det,
//     grad phi_0 = -(sum).
// 2D: det = e1 x e2 = 2 area (signed).
struct Geo {
  double g[4][3];
  double meas;         // |T|
  double signed_meas;  // signed |T|
};

__device__ __forceinline__ void element_geometry(const double4* __restrict__ X, int4 c, int dim, Geo& G) {
  double4 p0 = X[c.x], p1 = X[c.y], p2 = X[c.z];
  double e1x = p1.x - p0.x, e1y = p1.y - p0.y, e1z = p1.z - p0.z;
  double e2x = p2.x - p0.x, e2y = p2.y - p0.y, e2z = p2.z - p0.z;
  if (dim == 3) {
    double4 p3 = X[c.w];
    double e3x = p3.x - p0.x, e3y = p3.y - p0.y, e3z = p3.z - p0.z;
    // cofactor rows
    double c1x = e2y * e3z - e2z * e3y, c1y = e2z * e3x - e2x * e3z, c1z = e2x * e3y - e2y * e3x;
    double c2x = e3y * e1z - e3z * e1y, c2y = e3z * e1x - e3x * e1z, c2z = e3x * e1y - e3y * e1x;
    double c3x = e1y * e2z - e1z * e2y, c3y = e1z * e2x - e1x * e2z, c3z = e1x * e2y - e1y * e2x;
    double det = e1x * c1x + e1y * c1y + e1z * c1z;
    double inv = 1.0 / det;
    G.g[1][0] = c1x * inv; G.g[1][1] = c1y * inv; G.g[1][2] = c1z * inv;
    G.g[2][0] = c2x * inv; G.g[2][1] = c2y * inv; G.g[2][2] = c2z * inv;
    G.g[3][0] = c3x * inv; G.g[3][1] = c3y * inv; G.g[3][2] = c3z * inv;
    for (int d = 0; d < 3; ++d) G.g[0][d] = -(G.g[1][d] + G.g[2][d] + G.g[3][d]);
    G.signed_meas = det * (1.0 / 6.0);
    G.meas = fabs(G.signed_meas);
  } else {
    double det = e1x * e2y - e1y * e2x;
    double inv = 1.0 / det;
    G.g[1][0] = e2y * inv;  G.g[1][1] = -e2x * inv; G.g[1][2] = 0.0;
    G.g[2][0] = -e1y * inv; G.g[2][1] = e1x * inv;  G.g[2][2] = 0.0;
    G.g[0][0] = -(G.g[1][0] + G.g[2][0]);
    G.g[0][1] = -(G.g[1][1] + G.g[2][1]);
    G.g[0][2] = 0.0;
    G.g[3][0] = G.g[3][1] = G.g[3][2] = 0.0;
    G.signed_meas = 0.5 * det;
    G.meas = fabs(G.signed_meas);
  }
}

// ---- element-centric assembly ---------------------------------------------------
// Pass 1 (element_values_kernel): one thread per element computes, once, the
// element's off-diagonal stiffness entries S_ab = kappa |T| grad(phi_a).grad(phi_b)
// (a < b: 01 02 03 12 13 23 for tets, 01 02 12 for triangles) and its mass value
// m = mass_coeff |T| / 20 (tets) or / 12 (triangles; signed |T| with flags bit 0),
// stored as one record per element (8 doubles for tets, 4 for triangles).
// The element matrix is then K_ab = S_ab + m (a != b) and, because the
// stiffness rows of a P1 element sum to zero, K_aa = -sum_{b != a} S_ab + 2 m.
//
// Pass 2 (assemble_slices_kernel): a block owns a slice of 64 consecutive
// rows; each row walks the elements incident to it in ascending element
// order and adds the row of each element matrix into the row's entries.
// Every entry is a sum over its elements in ascending element order, and
// entry (i, j) and entry (j, i) add the same stored S_ab + m, so A is
// deterministic and bitwise symmetric without float atomics.
// The incidences are stored per slice, lane-interleaved ([k][64]: incidence k
// of the slice's 64 rows is one coalesced 512-byte read), each with the
// positions of the element's vertices in the row (one byte each), so no
// column search runs at assembly time.  The slice's rows are contiguous in
// the CSR value array: they are summed in LDS at their CSR offsets and
// written back by the block as one coalesced run.  Slices whose rows
// span more than kSliceSpan entries, or hold a row longer than 255 entries,
// sum in place in the CSR values with a column search (any row length).

template <int NPE>
struct EvLayout;
template <>
struct EvLayout<4> {
  static constexpr int kStride = 8;  // 6 S_ab, m, unused
  static constexpr int kM = 6;
  // slot of S_ab for corners a != b
  __device__ static int slot(int a, int b) {
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    return lo == 0 ? hi - 1 : (lo == 1 ? hi + 1 : 5);
  }
};
template <>
struct EvLayout<3> {
  static constexpr int kStride = 4;  // 3 S_ab, m
  static constexpr int kM = 3;
  __device__ static int slot(int a, int b) { return a + b - 1; }
};

// One thread per element; the block's records (consecutive elements) are
// staged in LDS and written back as one contiguous run, each store
// instruction covering 64 consecutive 16-byte pieces (a thread's own 64-byte
// record written directly would be four strided 16-byte pieces per lane).
template <int NPE>
__global__ void __launch_bounds__(256) element_values_kernel(int64_t ne, int signed_mass, double mass_coeff,
                                                             const int4* __restrict__ conn,
                                                             const double4* __restrict__ X,
                                                             const double* __restrict__ kappa,
                                                             double* __restrict__ ev) {
  constexpr int S = EvLayout<NPE>::kStride;
  constexpr int P = S + 1;  // padded LDS stride (doubles): conflict-free record writes
  __shared__ double rec[256 * P];
  // XCD-contiguous block order: neighbouring elements share vertices, and
  // each XCD's L2 then reads a contiguous range of the coordinates
  const int64_t e0 = (int64_t)xcd_block(blockIdx.x, gridDim.x) * 256;
  const int64_t e = e0 + threadIdx.x;
  if (e < ne) {
    // connectivity and coefficient are read once: nontemporal, so they do
    // not evict the coordinates the neighbouring elements gather
    typedef int i4 __attribute__((ext_vector_type(4)));
    const i4 cv = __builtin_nontemporal_load(reinterpret_cast<const i4*>(conn + e));
    const int4 c = make_int4(cv.x, cv.y, cv.z, cv.w);
    Geo G;
    element_geometry(X, c, NPE - 1, G);
    const double kv = __builtin_nontemporal_load(&kappa[e]) * G.meas;
    const double m = mass_coeff * (signed_mass ? G.signed_meas : G.meas) * (NPE == 4 ? 1.0 / 20.0 : 1.0 / 12.0);
    double* o = rec + threadIdx.x * P;
    if (NPE == 4) {
      int t = 0;
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = a + 1; b < 4; ++b)
          o[t++] = kv * (G.g[a][0] * G.g[b][0] + G.g[a][1] * G.g[b][1] + G.g[a][2] * G.g[b][2]);
      o[6] = m;
      o[7] = 0.0;
    } else {
      o[0] = kv * (G.g[0][0] * G.g[1][0] + G.g[0][1] * G.g[1][1]);
      o[1] = kv * (G.g[0][0] * G.g[2][0] + G.g[0][1] * G.g[2][1]);
      o[2] = kv * (G.g[1][0] * G.g[2][0] + G.g[1][1] * G.g[2][1]);
      o[3] = m;
    }
  }
  __syncthreads();
  const int64_t nrec = ne - e0 < 256 ? ne - e0 : 256;
  double2* out = reinterpret_cast<double2*>(ev + e0 * S);
  for (int i = threadIdx.x; i < nrec * (S / 2); i += 256) {
    const int r = i / (S / 2), part = i % (S / 2);
    out[i] = make_double2(rec[r * P + 2 * part], rec[r * P + 2 * part + 1]);
  }
}

// The NPE contributions of corner a of an element record r to its row:
// val[b] = S_ab + m (b != a), val[a] = 2 m - sum_{b != a} S_ab
template <int NPE>
__device__ __forceinline__ void corner_values(int a, const double* r, double val[4]) {
  const double m = r[EvLayout<NPE>::kM];
  double sum = 0.0;
#pragma unroll
  for (int b = 0; b < NPE; ++b) {
    if (b == a) continue;
    const double sab = r[EvLayout<NPE>::slot(a, b)];
    sum += sab;
    val[b] = sab + m;
  }
  val[a] = 2.0 * m - sum;
}

// element record of element e (8 doubles for tets, 4 for triangles)
template <int NPE>
__device__ __forceinline__ void load_record(const double* __restrict__ ev, int e, double r[8]) {
  const double4* p = reinterpret_cast<const double4*>(ev + (int64_t)e * EvLayout<NPE>::kStride);
  const double4 a = p[0];
  r[0] = a.x; r[1] = a.y; r[2] = a.z; r[3] = a.w;
  if (NPE == 4) {
    const double4 b = p[1];
    r[4] = b.x; r[5] = b.y; r[6] = b.z; r[7] = b.w;
  }
}

// The NPE contributions of corner a of element code (= 4 e + a) to its row:
// columns vid[] and values val[] (the diagonal one at index a)
template <int NPE>
__device__ __forceinline__ void corner_row(int code, const int4* __restrict__ conn, const double* __restrict__ ev,
                                           int vid[4], double val[4]) {
  const int e = code >> 2, a = code & 3;
  const int4 c = conn[e];
  vid[0] = c.x;
  vid[1] = c.y;
  vid[2] = c.z;
  vid[3] = c.w;
  double r[8];
  load_record<NPE>(ev, e, r);
  corner_values<NPE>(a, r, val);
}

constexpr int kSliceSpan = 2048;  // LDS doubles per slice (16 KiB)
constexpr int kIncGroup = 4;      // incidences in flight per lane in assemble_slices_kernel

// per slice of 64 rows: the most incidences of a row (0 for a slice that
// takes the in-place path) and whether the slice takes that path
__global__ void inc_slice_kernel(int64_t nv, const int32_t* __restrict__ v2e_ptr, const int32_t* __restrict__ rp,
                                 int32_t* __restrict__ kmax, int32_t* __restrict__ span, int32_t* __restrict__ is_long) {
  const int64_t sl = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nsl = (nv + 63) / 64;
  if (sl >= nsl) return;
  const int64_t r0 = sl * 64, r1 = r0 + 64 < nv ? r0 + 64 : nv;
  int km = 0, ml = 0;
  for (int64_t r = r0; r < r1; ++r) {
    km = max(km, v2e_ptr[r + 1] - v2e_ptr[r]);
    ml = max(ml, rp[r + 1] - rp[r]);
  }
  const int sp = rp[r1] - rp[r0];
  const int lg = (sp > kSliceSpan || ml > 255) ? 1 : 0;
  kmax[sl] = lg ? 0 : km;
  span[sl] = sp;
  is_long[sl] = lg;
}

// lane-interleaved incidences of the slices that sum in LDS: (element code,
// positions of the element's vertices in the row, one byte each); code -1
// pads a row's list to the slice's longest
template <int NPE>
__global__ void inc_fill_kernel(int64_t nv, const int32_t* __restrict__ v2e_ptr, const int32_t* __restrict__ v2e,
                                const int4* __restrict__ conn, const int32_t* __restrict__ rp,
                                const int32_t* __restrict__ ci, const int32_t* __restrict__ sptr,
                                int2* __restrict__ inc) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= nv) return;
  const int64_t sl = row >> 6;
  const int lane = (int)(row & 63);
  const int k0 = sptr[sl], kn = sptr[sl + 1] - k0;
  if (kn == 0) return;
  const int o = rp[row], len = rp[row + 1] - o;
  const int p0 = v2e_ptr[row], cnt = v2e_ptr[row + 1] - p0;
  for (int k = 0; k < kn; ++k) {
    int2 out = make_int2(-1, 0);
    if (k < cnt) {
      const int code = v2e[p0 + k];
      const int4 c = conn[code >> 2];
      const int vid[4] = {c.x, c.y, c.z, c.w};
      unsigned slots = 0;
#pragma unroll
      for (int b = 0; b < NPE; ++b) {
        int lo = 0, hi = len;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (ci[o + mid] < vid[b]) lo = mid + 1; else hi = mid;
        }
        slots |= (unsigned)lo << (8 * b);
      }
      out = make_int2(code, (int)slots);
    }
    inc[((int64_t)(k0 + k) << 6) + lane] = out;
  }
}

// incidence read once per assembly: nontemporal, so it does not evict the
// element records, which the rows of up to four slices read through L2
__device__ __forceinline__ int2 load_inc(const int2* p) {
  const long long v = __builtin_nontemporal_load(reinterpret_cast<const long long*>(p));
  return make_int2((int)(v & 0xffffffffll), (int)(v >> 32));
}

// 64-bit value of quad lane J, broadcast to the four lanes of each quad (DPP)
template <int J>
__device__ __forceinline__ double quad_bcast(double v) {
  constexpr int kCtrl = J | (J << 2) | (J << 4) | (J << 6);  // quad_perm [J, J, J, J]
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(b & 0xffffffffll), kCtrl, 0xf, 0xf, false);
  const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(b >> 32), kCtrl, 0xf, 0xf, false);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// rows of the slices that sum in LDS (see above).  A block of 256 threads
// owns a slice, four lanes (a quad) per row.  For incidence (e, a) lane t of
// the quad reads 16-byte chunk t of the record, so the quad reads the record
// as one contiguous 64-byte piece (a single lane reading it whole issues
// four scattered 16-byte reads, and the address unit, not HBM, then bounds
// the pass); lane t then takes the slot it needs from the lane holding it
// (ds_bpermute): S_at for t != a, m on lane a, and the quad shares the four
// values by DPP broadcasts.  Lane t adds S_at + m into the row's entry of
// column t in LDS, lane 0 sums the diagonal contribution
// 2 m - sum_{b != a} S_ab in a register.  Every entry still sums its
// elements in incidence order (the columns of one incidence are distinct),
// with the values and order of corner_values.  With four lanes a row a block
// holds one slice, so a CU has few slices in flight and the records they
// share stay in L2 between their reads.
template <int NPE, int U>
__global__ void __launch_bounds__(256) assemble_slices_kernel(int64_t nv, const int32_t* __restrict__ sptr,
                                                              const int32_t* __restrict__ is_long,
                                                              const int2* __restrict__ inc,
                                                              const double* __restrict__ ev,
                                                              const int32_t* __restrict__ rp,
                                                              double* __restrict__ val) {
  extern __shared__ double acc[];
  const int t = threadIdx.x & 3;
  const int row = threadIdx.x >> 2;  // row of the slice
  const int64_t nsl = (nv + 63) / 64;
  const XcdRows R = xcd_rows(nsl, 1);
  for (int64_t sl = R.first; sl < R.hi; sl += R.step) {
    // slices with a wide span or a row over 255 entries sum in place elsewhere
    if (is_long[sl]) continue;
    const int64_t r0 = sl * 64, r1 = r0 + 64 < nv ? r0 + 64 : nv;
    const int64_t o0 = rp[r0];
    const int span = rp[r1] - (int)o0;
    // every lane runs the slice's incidence count (rows past the last one
    // hold padding), so all lanes of a quad take part in the broadcasts
    const int kn = sptr[sl + 1] - sptr[sl];
    const int base = r0 + row < nv ? rp[r0 + row] - (int)o0 : 0;
    for (int i = threadIdx.x; i < span; i += 256) acc[i] = 0.0;
    __syncthreads();
    const int2* q = inc + ((int64_t)sptr[sl] << 6) + row;
    double diag = 0.0;
    int dslot = -1;
    // software pipeline over groups of U incidences: the next group's
    // incidences load while this group's records are in flight; a padding
    // incidence reads element 0's record and adds nothing
    const int ng = (kn + U - 1) / U;
    int2 nx[U];
    auto inc_at = [&](int k) { return k < kn ? load_inc(q + ((int64_t)k << 6)) : make_int2(-1, 0); };
#pragma unroll
    for (int u = 0; u < U; ++u) nx[u] = inc_at(u);
    for (int g = 0; g < ng; ++g) {
      int2 ic[U];
      double2 ch[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        ic[u] = nx[u];
        const double* rec = ev + (int64_t)(ic[u].x >= 0 ? ic[u].x >> 2 : 0) * EvLayout<NPE>::kStride;
        ch[u] = reinterpret_cast<const double2*>(rec)[t % (EvLayout<NPE>::kStride / 2)];
      }
      const int kk = (g + 1) * U;
#pragma unroll
      for (int u = 0; u < U; ++u) nx[u] = inc_at(kk + u);
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int a = ic[u].x & 3;
        // lane t takes S_at from the lane holding its slot (the corner's own
        // lane, and lane 3 of a triangle, take m)
        const int sl_t = (t == a || t >= NPE) ? EvLayout<NPE>::kM : EvLayout<NPE>::slot(a, t);
        const int src = (((int)threadIdx.x & 63 & ~3) + (sl_t >> 1)) << 2;
        const long long c0 = __double_as_longlong(ch[u].x), c1 = __double_as_longlong(ch[u].y);
        const unsigned lo0 = __builtin_amdgcn_ds_bpermute(src, (int)(c0 & 0xffffffffll));
        const unsigned hi0 = __builtin_amdgcn_ds_bpermute(src, (int)(c0 >> 32));
        const unsigned lo1 = __builtin_amdgcn_ds_bpermute(src, (int)(c1 & 0xffffffffll));
        const unsigned hi1 = __builtin_amdgcn_ds_bpermute(src, (int)(c1 >> 32));
        const double x = __longlong_as_double((long long)((sl_t & 1) ? (((unsigned long long)hi1 << 32) | lo1)
                                                                     : (((unsigned long long)hi0 << 32) | lo0)));
        // sum_{b != a} S_ab in ascending b, as corner_values adds it; m from lane a
        double sum = 0.0;
        const double y0 = quad_bcast<0>(x), y1 = quad_bcast<1>(x), y2 = quad_bcast<2>(x);
        const double y3 = quad_bcast<3>(x);
        if (a != 0) sum += y0;
        if (a != 1) sum += y1;
        if (a != 2) sum += y2;
        if (NPE == 4 && a != 3) sum += y3;
        const double m = a == 0 ? y0 : (a == 1 ? y1 : (a == 2 ? y2 : y3));
        if (ic[u].x >= 0) {
          if (t < NPE && t != a) acc[base + ((ic[u].y >> (8 * t)) & 255)] += x + m;
          if (t == 0) {
            diag += 2.0 * m - sum;
            dslot = (ic[u].y >> (8 * a)) & 255;
          }
        }
      }
    }
    if (t == 0 && dslot >= 0) acc[base + dslot] = diag;
    __syncthreads();
    for (int i = threadIdx.x; i < span; i += 256) __builtin_nontemporal_store(acc[i], val + o0 + i);
    __syncthreads();
  }
}

__device__ __forceinline__ int find_col(const int32_t* __restrict__ cols, int len, int c) {
  int lo = 0, hi = len;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (cols[mid] < c) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// rows of the slices listed in long_sl: sums in place in the CSR values,
// columns found by binary search (any row length)
template <int NPE>
__global__ void __launch_bounds__(64) assemble_rows_long_kernel(int64_t nv, int64_t nlong,
                                                                const int32_t* __restrict__ long_sl,
                                                                const int32_t* __restrict__ v2e_ptr,
                                                                const int32_t* __restrict__ v2e,
                                                                const int4* __restrict__ conn,
                                                                const double* __restrict__ ev,
                                                                const int32_t* __restrict__ rp,
                                                                const int32_t* __restrict__ ci,
                                                                double* __restrict__ val) {
  if ((int64_t)blockIdx.x >= nlong) return;
  const int64_t row = (int64_t)long_sl[blockIdx.x] * 64 + threadIdx.x;
  if (row >= nv) return;
  const int o = rp[row];
  const int len = rp[row + 1] - o;
  const int32_t* cols = ci + o;
  double* acc = val + o;
  for (int t = 0; t < len; ++t) acc[t] = 0.0;
  for (int p = v2e_ptr[row]; p < v2e_ptr[row + 1]; ++p) {
    int vid[4];
    double v[4];
    corner_row<NPE>(v2e[p], conn, ev, vid, v);
#pragma unroll
    for (int b = 0; b < NPE; ++b) acc[find_col(cols, len, vid[b])] += v[b];
  }
}

// dL/dkappa_e = -|T_e| grad(a)_e . grad(u)_e, written in user element order.
__global__ void element_gradient_kernel(int64_t ne, int dim, int npe, const int4* __restrict__ conn,
                                        const double4* __restrict__ X, const int32_t* __restrict__ eperm,
                                        const double* __restrict__ a, const double* __restrict__ u,
                                        double* __restrict__ out) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  int4 c = conn[e];
  Geo G;
  element_geometry(X, c, dim, G);
  const int v[4] = {c.x, c.y, c.z, c.w};
  double ga[3] = {0, 0, 0}, gu[3] = {0, 0, 0};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (k < npe) {
      double av = a[v[k]], uv = u[v[k]];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        ga[d] += av * G.g[k][d];
        gu[d] += uv * G.g[k][d];
      }
    }
  }
  out[eperm[e]] = -G.meas * (ga[0] * gu[0] + ga[1] * gu[1] + ga[2] * gu[2]);
}

__global__ void kappa_gather_kernel(int64_t ne, const int32_t* __restrict__ eperm, const double* __restrict__ ku,
                                    double* __restrict__ k) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < ne) k[e] = ku[eperm[e]];
}

__global__ void gather_kernel(int64_t n, int k, const int32_t* __restrict__ perm, const double* __restrict__ in,
                              double* __restrict__ out) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * k) return;
  int64_t i = t / k, c = t - i * k;
  out[t] = in[(int64_t)perm[i] * k + c];
}

__global__ void scatter_kernel(int64_t n, int k, const int32_t* __restrict__ perm, const double* __restrict__ in,
                               double* __restrict__ out) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * k) return;
  int64_t i = t / k, c = t - i * k;
  out[(int64_t)perm[i] * k + c] = in[t];
}

// bit 0: a vertex index outside [0, nv); bit 1: NaN coefficient; bit 2: NaN coordinate
__global__ void mesh_check_kernel(int64_t nconn, const int32_t* __restrict__ conn, int64_t nv, int64_t ne,
                                  const double* __restrict__ kappa, int64_t nx, const double* __restrict__ xyz,
                                  int* __restrict__ flags) {
  const int64_t st = (int64_t)gridDim.x * blockDim.x;
  int f = 0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nconn; t += st) {
    const int c = conn[t];
    if (c < 0 || c >= nv) f |= 1;
  }
  if (kappa)
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < ne; t += st)
      if (!(kappa[t] == kappa[t])) f |= 2;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nx; t += st)
    if (!(xyz[t] == xyz[t])) f |= 4;
  if (f) atomicOr(flags, f);
}

// per-block (min x, min y, min z, max x, max y, max z) of the vertex coordinates
__global__ void __launch_bounds__(256) bbox_kernel(int64_t nv, const double* __restrict__ xyz,
                                                   double* __restrict__ part) {
  __shared__ double s_v[6][256];
  double v[6] = {1e300, 1e300, 1e300, -1e300, -1e300, -1e300};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += (int64_t)gridDim.x * blockDim.x)
    for (int d = 0; d < 3; ++d) {
      const double c = xyz[3 * i + d];
      v[d] = fmin(v[d], c);
      v[3 + d] = fmax(v[3 + d], c);
    }
  for (int j = 0; j < 6; ++j) s_v[j][threadIdx.x] = v[j];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int j = 0; j < 6; ++j)
        s_v[j][threadIdx.x] = j < 3 ? fmin(s_v[j][threadIdx.x], s_v[j][threadIdx.x + w])
                                    : fmax(s_v[j][threadIdx.x], s_v[j][threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = s_v[threadIdx.x][0];
}

}  // namespace

int check_mesh_device(int64_t nconn, const int32_t* conn, int64_t nv, int64_t ne, const double* kappa,
                      int64_t nx, const double* xyz, hipStream_t s) {
  DevBuf<int> flags(1);
  dev_fill_i32(flags.get(), 0, 1, s);
  const int64_t m = std::max(nconn, std::max(ne, nx));
  mesh_check_kernel<<<(unsigned)std::min<int64_t>(std::max<int64_t>((m + 255) / 256, 1), 4096), 256, 0, s>>>(
      nconn, conn, nv, ne, kappa, nx, xyz, flags.get());
  HIP_LAUNCH_CHECK();
  int f = 0;
  HIP_CHECK(hipMemcpyAsync(&f, flags.get(), sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  return f;
}

void plan_assembly_slices(Handle& h);

// radix-sort passes cover only the bits that keys below `bound` use (the
// sort is stable, so the order is the same as over all 32 bits)
static int key_bits(int64_t bound) {
  int b = 1;
  while (b < 32 && ((int64_t)1 << b) < bound) ++b;
  return b;
}

void build_mesh_device(Handle& h, const double* xyz, const int32_t* conn, const double* kappa) {
  hipStream_t s = h.stream;
  const int64_t nv = h.nv, ne = h.ne;
  const int npe = h.dim + 1;
  if (nv <= 0 || ne <= 0) fail(FEM_ERR_ARG, "mesh has no vertices or no elements");
  if (4 * ne >= (int64_t)1 << 31 || nv >= (int64_t)1 << 31)
    fail(FEM_ERR_CAPACITY, "mesh too large for 32-bit element/vertex indices");

  DevBuf<double> xyz_u((size_t)(3 * nv));
  DevBuf<int32_t> conn_u((size_t)(npe * ne));
  DevBuf<double> kappa_u(ne);
  {
    const HostPiece pc[3] = {{xyz_u.get(), xyz, sizeof(double) * (size_t)(3 * nv)},
                             {conn_u.get(), conn, sizeof(int32_t) * (size_t)(npe * ne)},
                             {kappa_u.get(), kappa, kappa ? sizeof(double) * (size_t)ne : 0}};
    upload_pieces(pc, 3, s);
  }
  if (!kappa) dev_fill_f64(kappa_u.get(), 1.0, (size_t)ne, s);
  const int bad = check_mesh_device(npe * ne, conn_u.get(), nv, ne, kappa ? kappa_u.get() : nullptr, 3 * nv,
                                    xyz_u.get(), s);
  if (bad & 1) fail(FEM_ERR_ARG, "element vertex index out of range");
  if (bad & 2) fail(FEM_ERR_NUMERIC, "NaN in element coefficient");
  if (bad & 4) fail(FEM_ERR_NUMERIC, "NaN in vertex coordinates");

  // bounding box: per-block min/max on the device, the last fold on the host
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  {
    const unsigned nb = (unsigned)std::min<int64_t>((nv + 255) / 256, 1024);
    DevBuf<double> part(6 * (size_t)nb);
    bbox_kernel<<<nb, 256, 0, s>>>(nv, xyz_u.get(), part.get());
    HIP_LAUNCH_CHECK();
    std::vector<double> hp(6 * (size_t)nb);
    part.download(hp.data(), hp.size(), s);
    HIP_CHECK(hipStreamSynchronize(s));
    for (unsigned b = 0; b < nb; ++b)
      for (int d = 0; d < 3; ++d) {
        lo[d] = std::min(lo[d], hp[6 * b + d]);
        hi[d] = std::max(hi[d], hp[6 * b + 3 + d]);
      }
  }
  double sc[3];
  for (int d = 0; d < 3; ++d) sc[d] = (hi[d] > lo[d]) ? 1.0 / (hi[d] - lo[d]) : 0.0;

  // vertex order
  {
    DevBuf<uint64_t> key(nv), key2(nv);
    DevBuf<int32_t> idx(nv);
    h.perm.resize(nv);
    h.iperm.resize(nv);
    morton_kernel<<<grid_for(nv, 256), 256, 0, s>>>(nv, h.dim, xyz_u.get(), lo[0], lo[1], lo[2], sc[0], sc[1],
                                                    sc[2], key.get(), idx.get());
    HIP_LAUNCH_CHECK();
    dev_sort_pairs_u64_i32(key.get(), key2.get(), idx.get(), h.perm.get(), (size_t)nv, 64, s);
    invert_perm_kernel<<<grid_for(nv, 256), 256, 0, s>>>(nv, h.perm.get(), h.iperm.get());
    HIP_LAUNCH_CHECK();
    h.coords.resize(4 * nv);
    coords_kernel<<<grid_for(nv, 256), 256, 0, s>>>(nv, xyz_u.get(), h.perm.get(), (double4*)h.coords.get());
    HIP_LAUNCH_CHECK();
  }
  // element order: ascending lowest internal vertex (stable)
  {
    DevBuf<int32_t> tmp(4 * ne);
    DevBuf<uint32_t> key(ne), key2(ne);
    DevBuf<int32_t> idx(ne);
    relabel_elems_kernel<<<grid_for(ne, 256), 256, 0, s>>>(ne, npe, conn_u.get(), h.iperm.get(),
                                                           (int4*)tmp.get(), key.get(), idx.get());
    HIP_LAUNCH_CHECK();
    h.eperm.resize(ne);
    dev_sort_pairs_u32_i32(key.get(), key2.get(), idx.get(), h.eperm.get(), (size_t)ne, key_bits(nv), s);
    h.conn.resize(4 * ne);
    h.kappa.resize(ne);
    permute_elems_kernel<<<grid_for(ne, 256), 256, 0, s>>>(ne, h.eperm.get(), (const int4*)tmp.get(),
                                                           kappa_u.get(), (int4*)h.conn.get(), h.kappa.get());
    HIP_LAUNCH_CHECK();
  }
  // vertex -> element incidence
  {
    int64_t m = npe * ne;
    DevBuf<uint32_t> key(m), key2(m);
    DevBuf<int32_t> val(m);
    v2e_keys_kernel<<<grid_for(ne, 256), 256, 0, s>>>(ne, npe, (const int4*)h.conn.get(), key.get(), val.get());
    HIP_LAUNCH_CHECK();
    h.v2e.resize(m);
    dev_sort_pairs_u32_i32(key.get(), key2.get(), val.get(), h.v2e.get(), (size_t)m, key_bits(nv), s);
    h.v2e_ptr.resize(nv + 1);
    lower_bound_kernel<<<grid_for(nv + 1, 256), 256, 0, s>>>(nv, key2.get(), m, h.v2e_ptr.get());
    HIP_LAUNCH_CHECK();
  }
  // pattern
  {
    DevBuf<int32_t> cnt(nv);
    // sorted rows of the count pass, kRowCap slots each (lane-interleaved per 64 rows)
    DevBuf<int32_t> stash((size_t)((nv + kRowBlock - 1) / kRowBlock) * kRowBlock * kRowCap);
    pattern_rows_kernel<false><<<grid_for(nv, kRowBlock), kRowBlock, 0, s>>>(
        nv, npe, h.v2e_ptr.get(), h.v2e.get(), (const int4*)h.conn.get(), cnt.get(), nullptr, nullptr, stash.get());
    HIP_LAUNCH_CHECK();
    // long rows
    DevBuf<int32_t> long_rows(nv);
    DevBuf<int64_t> cap(nv + 1);
    DevBuf<unsigned> nlong_d(1);
    HIP_CHECK(hipMemsetAsync(nlong_d.get(), 0, sizeof(unsigned), s));
    collect_long_rows_kernel<<<grid_for(nv, 256), 256, 0, s>>>(nv, cnt.get(), h.v2e_ptr.get(), npe,
                                                               long_rows.get(), cap.get(), nlong_d.get());
    HIP_LAUNCH_CHECK();
    unsigned nlong = 0;
    HIP_CHECK(hipMemcpyAsync(&nlong, nlong_d.get(), sizeof(unsigned), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    DevBuf<int64_t> off;
    DevBuf<int> scratch;
    if (nlong) {
      // deterministic processing order of long rows does not matter: each is independent
      off.resize(nlong + 1);
      int64_t tot = dev_exclusive_scan_i64(cap.get(), off.get(), nlong, s);
      scratch.resize(tot);
      pattern_rows_global_kernel<false><<<grid_for(nlong, 64), 64, 0, s>>>(
          nlong, long_rows.get(), off.get(), npe, h.v2e_ptr.get(), h.v2e.get(), (const int4*)h.conn.get(),
          scratch.get(), cnt.get(), nullptr, nullptr);
      HIP_LAUNCH_CHECK();
    }
    h.A.rp.resize(nv + 1);
    int64_t nnz = dev_exclusive_scan_i32(cnt.get(), h.A.rp.get(), nv, s);
    if (nnz >= ((int64_t)1 << 31)) fail(FEM_ERR_CAPACITY, "matrix has 2^31 or more nonzeros");
    h.A.n = h.A.m = nv;
    h.A.nnz = nnz;
    h.A.ci.resize(nnz);
    // without long rows every row is sorted in LDS by the fill pass, which
    // then also writes the assembly incidences (no separate search pass)
    const bool with_inc = nlong == 0;
    if (with_inc) plan_assembly_slices(h);
    pattern_rows_kernel<true><<<grid_for(nv, kRowBlock), kRowBlock, 0, s>>>(
        nv, npe, h.v2e_ptr.get(), h.v2e.get(), (const int4*)h.conn.get(), cnt.get(), h.A.rp.get(), h.A.ci.get(),
        stash.get(), with_inc ? h.asm_sptr.get() : nullptr, with_inc ? h.asm_inc.get() : nullptr);
    HIP_LAUNCH_CHECK();
    if (with_inc) h.asm_ready = true;
    if (nlong) {
      pattern_rows_global_kernel<true><<<grid_for(nlong, 64), 64, 0, s>>>(
          nlong, long_rows.get(), off.get(), npe, h.v2e_ptr.get(), h.v2e.get(), (const int4*)h.conn.get(),
          scratch.get(), cnt.get(), h.A.rp.get(), h.A.ci.get());
      HIP_LAUNCH_CHECK();
    }
    HIP_CHECK(hipStreamSynchronize(s));
  }
  h.A.v64.resize(h.A.nnz);
  h.has_mesh = true;
  assemble_values(h);
}

// Slice layout of the incidences for assemble_slices_kernel (once per
// pattern; re-assembly with new coefficients reuses it): slice offsets, the
// slices that sum in place, and the incidence array filled with padding.
void plan_assembly_slices(Handle& h) {
  hipStream_t s = h.stream;
  const int64_t nv = h.nv, nsl = (nv + 63) / 64;
  DevBuf<int32_t> kmax(nsl), span(nsl), lg(nsl);
  inc_slice_kernel<<<grid_for(nsl, 256), 256, 0, s>>>(nv, h.v2e_ptr.get(), h.A.rp.get(), kmax.get(), span.get(),
                                                      lg.get());
  HIP_LAUNCH_CHECK();
  h.asm_sptr.resize(nsl + 1);
  const int64_t tot = dev_exclusive_scan_i32(kmax.get(), h.asm_sptr.get(), (size_t)nsl, s);
  if (tot * 64 >= ((int64_t)1 << 31)) fail(FEM_ERR_CAPACITY, "too many element incidences for the assembly layout");
  std::vector<int32_t> hs(nsl), hl(nsl);
  span.download(hs.data(), (size_t)nsl, s);
  lg.download(hl.data(), (size_t)nsl, s);
  HIP_CHECK(hipStreamSynchronize(s));
  int smax = 1;
  std::vector<int32_t> longs;
  for (int64_t i = 0; i < nsl; ++i) {
    if (hl[i]) longs.push_back((int32_t)i);
    else smax = std::max(smax, hs[i]);
  }
  h.asm_span = smax;
  h.asm_nlong = (int64_t)longs.size();
  h.asm_long.resize(longs.empty() ? 1 : longs.size());
  if (!longs.empty()) h.asm_long.upload(longs.data(), longs.size(), s);
  h.asm_islong = std::move(lg);
  h.asm_inc.resize(tot > 0 ? (size_t)tot * 64 : 1);
  dev_fill_i32(reinterpret_cast<int32_t*>(h.asm_inc.get()), -1, 2 * h.asm_inc.size(), s);
}

// The layout and the incidences of a pattern built elsewhere (or with long rows).
void build_assembly_slices(Handle& h) {
  hipStream_t s = h.stream;
  const int64_t nv = h.nv;
  plan_assembly_slices(h);
  const int4* conn = (const int4*)h.conn.get();
  if (h.dim == 3)
    inc_fill_kernel<4><<<grid_for(nv, 256), 256, 0, s>>>(nv, h.v2e_ptr.get(), h.v2e.get(), conn, h.A.rp.get(),
                                                         h.A.ci.get(), h.asm_sptr.get(), h.asm_inc.get());
  else
    inc_fill_kernel<3><<<grid_for(nv, 256), 256, 0, s>>>(nv, h.v2e_ptr.get(), h.v2e.get(), conn, h.A.rp.get(),
                                                         h.A.ci.get(), h.asm_sptr.get(), h.asm_inc.get());
  HIP_LAUNCH_CHECK();
  HIP_CHECK(hipStreamSynchronize(s));
  h.asm_ready = true;
}

void assemble_values(Handle& h) {
  hipStream_t s = h.stream;
  h.A.nslices = 0;  // the sliced-ELL copy is rebuilt with the next hierarchy
  const int npe = h.dim + 1;
  const int stride = npe == 4 ? EvLayout<4>::kStride : EvLayout<3>::kStride;
  h.asm_ev.reserve((size_t)h.ne * stride);
  if (!h.asm_ready) build_assembly_slices(h);
  const int sm = (h.flags & 1) ? 1 : 0;
  const int4* conn = (const int4*)h.conn.get();
  double* ev = h.asm_ev.get();
  const int32_t* rp = h.A.rp.get();
  const int32_t* ci = h.A.ci.get();
  double* val = h.A.v64.get();
  // algorithmic bytes: connectivity, coefficient, coordinates, incidence, CSR columns and values
  ProfScope ps(h, "assembly", (double)h.ne * (16 + 8) + (double)h.nv * 32 + (double)h.A.nnz * 12 +
                                  (double)h.ne * npe * 4);
  const int64_t nsl = (h.nv + 63) / 64;
  const unsigned ge = grid_for(h.ne, 256);
  const unsigned gs = (unsigned)nsl;
  const size_t lds = sizeof(double) * (size_t)h.asm_span;
  if (npe == 4) {
    element_values_kernel<4><<<ge, 256, 0, s>>>(h.ne, sm, h.mass_coeff, conn, (const double4*)h.coords.get(),
                                                h.kappa.get(), ev);
    HIP_LAUNCH_CHECK();
    assemble_slices_kernel<4, kIncGroup><<<gs, 256, lds, s>>>(h.nv, h.asm_sptr.get(), h.asm_islong.get(),
                                                              h.asm_inc.get(), ev, rp, val);
    HIP_LAUNCH_CHECK();
    if (h.asm_nlong)
      assemble_rows_long_kernel<4><<<(unsigned)h.asm_nlong, 64, 0, s>>>(h.nv, h.asm_nlong, h.asm_long.get(),
                                                                         h.v2e_ptr.get(), h.v2e.get(), conn, ev, rp,
                                                                         ci, val);
  } else {
    element_values_kernel<3><<<ge, 256, 0, s>>>(h.ne, sm, h.mass_coeff, conn, (const double4*)h.coords.get(),
                                                h.kappa.get(), ev);
    HIP_LAUNCH_CHECK();
    assemble_slices_kernel<3, kIncGroup><<<gs, 256, lds, s>>>(h.nv, h.asm_sptr.get(), h.asm_islong.get(),
                                                              h.asm_inc.get(), ev, rp, val);
    HIP_LAUNCH_CHECK();
    if (h.asm_nlong)
      assemble_rows_long_kernel<3><<<(unsigned)h.asm_nlong, 64, 0, s>>>(h.nv, h.asm_nlong, h.asm_long.get(),
                                                                         h.v2e_ptr.get(), h.v2e.get(), conn, ev, rp,
                                                                         ci, val);
  }
  HIP_LAUNCH_CHECK();
}

void element_gradient(Handle& h, const double* a_int, const double* u_int, double* g_user) {
  ProfScope ps(h, "element_grad", (double)h.ne * (16 + 4 + 8) + (double)h.nv * (32 + 16));
  element_gradient_kernel<<<grid_for(h.ne, 256), 256, 0, h.stream>>>(
      h.ne, h.dim, h.dim + 1, (const int4*)h.conn.get(), (const double4*)h.coords.get(), h.eperm.get(), a_int,
      u_int, g_user);
  HIP_LAUNCH_CHECK();
}

void set_kappa_user(Handle& h, const double* kappa_user_dev) {
  kappa_gather_kernel<<<grid_for(h.ne, 256), 256, 0, h.stream>>>(h.ne, h.eperm.get(), kappa_user_dev, h.kappa.get());
  HIP_LAUNCH_CHECK();
}

void gather_user_to_int(Handle& h, const double* u, double* i, int k) {
  int64_t n = h.A.n;
  gather_kernel<<<grid_for(n * k, 256), 256, 0, h.stream>>>(n, k, h.perm.get(), u, i);
  HIP_LAUNCH_CHECK();
}

void scatter_int_to_user(Handle& h, const double* i, double* u, int k) {
  int64_t n = h.A.n;
  scatter_kernel<<<grid_for(n * k, 256), 256, 0, h.stream>>>(n, k, h.perm.get(), i, u);
  HIP_LAUNCH_CHECK();
}

}  // namespace fem
