// Internal data structures shared by the translation units of libfem_mi355x.
//
// Device layout (all in the library's internal vertex numbering, a Morton
// order of the vertex coordinates, so that rows that are close in the mesh are
// close in memory):
//   coords   : nv x double4 (x, y, z, 0)                         32 B / vertex
//   conn     : ne x int4 (tets) / int4 with w = -1 (triangles)    16 B / element
//   kappa    : ne doubles (element order sorted by lowest vertex)
//   v2e      : CSR vertex -> (element*4 + local corner), ascending element
//   A        : CSR int32 row_ptr / int32 cols (ascending) / FP64 values,
//              plus an FP32 copy of every level in mixed precision.
#pragma once

#include "../../include/fem_mi355x.h"
#include "common.hpp"

#include <map>
#include <memory>
#include <string>
#include <vector>

namespace fem {

constexpr int kSellWindows = 8;
constexpr int kSellWinBits = 13;

// Compressed sparse rows on the device.
struct Csr {
  int64_t n = 0;    // rows
  int64_t m = 0;    // columns
  int64_t nnz = 0;
  DevBuf<int32_t> rp;  // n + 1
  DevBuf<int32_t> ci;  // nnz
  DevBuf<double> v64;  // nnz
  DevBuf<float> v32;   // nnz (mixed precision only)
  // Sliced ELL copy for the solve phase: slices of 64 / sell_g consecutive
  // rows (one wavefront, sell_g lanes per row), stored column-major inside the
  // slice: entry j of the slice's row q sits at step j / sell_g of lane
  // q * sell_g + j % sell_g.  Slices are padded to their longest row (padding:
  // value 0, column = the row's last column).
  // Columns are kept twice: as int32 (sc) and as 16-bit codes (sc16) for the
  // compact slices, those whose columns fall in at most kSellWindows aligned
  // windows of 2^kSellWinBits columns.  swin holds each slice's window bases
  // (column offsets, multiples of 2^kSellWinBits; first entry -1 for a slice
  // that is not compact); a code is (window << kSellWinBits) | (col - base).
  int64_t nslices = 0;
  int sell_g = 1;       // lanes per row (1, 2, 4, 8, 16 or 32)
  DevBuf<int32_t> sp;   // nslices + 1 element offsets (multiples of 64)
  DevBuf<int32_t> sc;   // padded columns
  DevBuf<uint16_t> sc16;  // padded 16-bit column codes (compact slices)
  DevBuf<int32_t> swin;   // kSellWindows bases per slice
  int64_t ncompact = 0;     // number of compact slices
  int64_t nnz_compact = 0;  // stored nonzeros in compact slices
  DevBuf<int32_t> srow;     // row at each sliced position (empty: position = row)
  DevBuf<double> s64;   // padded FP64 values
  DevBuf<float> s32;    // padded FP32 values (mixed precision only)
  DevBuf<uint16_t> sh;  // padded FP16 values (bit patterns; precision 2, level 0 only)
  // Diagonal kept outside the slices (the fine operator): the slices hold the
  // off-diagonal entries and a row's diagonal term a_ii x_i is added from
  // these (empty: the diagonal is in the slices)
  DevBuf<double> sd64;
  DevBuf<float> sd32;
  int64_t nnz_sdiag = 0;  // diagonal entries held in sd64 / sd32
  // LDS staging of the gathered vector (one lane per row, row order): block b
  // of lds_bs consecutive slices first copies x at its distinct columns
  // lcols[lcs[b] .. lcs[b+1]) (ascending) into LDS; lc holds each stored
  // entry's 16-bit index into that list (same layout as sc).  lds_bs == 0: none.
  int lds_bs = 0;
  int lds_maxc = 0;            // largest distinct-column count of a block
  DevBuf<uint16_t> lc;
  DevBuf<int32_t> lcs, lcols;
};

struct Level {
  Csr A;
  DevBuf<double> dinv64;
  DevBuf<float> dinv32;
  Csr P;                  // n x nc prolongator
  Csr R;                  // nc x n restriction = P^T
  DevBuf<int32_t> agg;    // n, aggregate of each row
  int64_t nc = 0;
  double rho = 2.0;       // spectral radius estimate of D^-1 A
  // V-cycle work vectors (one RHS column block of width `wk`)
  int wk = 0;
  DevBuf<double> x64, t64, b64, r64, d64;
  DevBuf<float> x32, t32, b32, r32, d32;
};

struct Hierarchy {
  std::vector<std::unique_ptr<Level>> levels;  // levels[0].A aliases nothing: its own copy
  int coarse = -1;                             // index of the coarsest level
  int64_t nc = 0;                              // coarsest size
  DevBuf<double> cinv64;                       // nc x nc dense FP64 inverse (host Cholesky at setup),
                                               // applied in FP64 in every precision mode
  double grid_complexity = 0, operator_complexity = 0;
  double setup_ms = 0;
  int precision = 0;
  bool built = false;
  // precision 2: the level-0 smoother works on y = D^1/2 x with the operator
  // D^-1/2 A D^-1/2 = diag(ah_diag) + (off-diagonals in FP16, `ah`); the
  // rounding of each row's off-diagonals is moved into its diagonal so that
  // the operator keeps the row sums of A.  ph = D^1/2 P, rh = ph^T.
  Csr ah;
  DevBuf<float> ah_diag;  // 1 + eps_i
  DevBuf<float> ah_s;     // 1 / sqrt(a_ii)
  Csr ph, rh;
};

struct KernelStat {
  int64_t launches = 0;
  double total_ms = 0;
  double bytes = 0;
};

struct PendingTiming {
  std::string fam;
  hipEvent_t a, b;
  double bytes;
};

struct Handle {
  std::string err;
  fem_params prm{};
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int device = 0;

  // mesh sizes (the arrays themselves live on the device only)
  int dim = 3;
  int64_t nv = 0, ne = 0;
  double mass_coeff = 1.0;
  int flags = 0;

  // device mesh, internal numbering
  DevBuf<double> coords;     // nv x 4
  DevBuf<int32_t> conn;      // ne x 4
  DevBuf<double> kappa;      // ne
  DevBuf<int32_t> perm;      // internal -> user vertex
  DevBuf<int32_t> iperm;     // user -> internal vertex
  DevBuf<int32_t> eperm;     // internal -> user element
  DevBuf<int32_t> v2e_ptr;   // nv + 1
  DevBuf<int32_t> v2e;       // 4 ne (or 3 ne)
  bool has_mesh = false;
  bool matrix_from_file = false;

  // element-centric assembly: per-element records (off-diagonal stiffness
  // entries and the mass value, 8 doubles per tet / 4 per triangle), kept
  // for re-assembly; the incidences of each 64-row slice, lane-interleaved,
  // with the vertices' positions in the row (asm_inc, slice offsets in
  // 64-entry units in asm_sptr); the largest LDS span of a slice; the slices
  // that sum in place in the CSR values instead (asm_long)
  DevBuf<double> asm_ev;
  DevBuf<int2> asm_inc;
  DevBuf<int32_t> asm_sptr;
  DevBuf<int32_t> asm_long;
  DevBuf<int32_t> asm_islong;  // per slice: 1 when it sums in place
  int64_t asm_nlong = 0;
  int asm_span = 0;
  bool asm_ready = false;

  // level-0 operator (FP64, internal numbering)
  Csr A;
  std::vector<int64_t> mesh_pattern_rp;  // user-numbered mesh pattern (host)
  std::vector<int32_t> mesh_pattern_ci;

  Hierarchy H;

  // Krylov work (FP64), capacity for `kcap` right-hand sides
  int kcap = 0;
  DevBuf<double> kx, kr, kp, kq, kz, kb;
  DevBuf<double> partials;        // reduction partials
  DevBuf<double> scal;            // device scalars
  double* h_scal = nullptr;       // pinned mirror of scal
  double* h_act = nullptr;        // per-column convergence flags, host memory mapped
  double* d_act = nullptr;        //   for the device (written by the |r|^2 finish)
  DevBuf<double> ubuf, ubuf2;     // user-order staging (nv * k)

  // captured PCG iteration (launchMode 1), valid while gen == graph_gen
  uint64_t gen = 1;             // bumped when buffers, hierarchy or parameters change
  hipGraphExec_t iter_graph = nullptr;
  uint64_t graph_gen = 0;
  int graph_k = 0;
  double graph_tol2 = 0.0;

  bool profiling = false;
  std::map<std::string, KernelStat> stats;
  std::vector<PendingTiming> pending;   // recorded, not yet read back
  std::vector<hipEvent_t> ev_pool;      // free events
};

// ---- primitives (primitives.hip) ----
void dev_sort_pairs_u64_i32(const uint64_t* kin, uint64_t* kout, const int32_t* vin, int32_t* vout,
                            size_t n, int end_bit, hipStream_t s);
void dev_sort_pairs_u32_i32(const uint32_t* kin, uint32_t* kout, const int32_t* vin, int32_t* vout,
                            size_t n, int end_bit, hipStream_t s);
// out[0] = 0, out[i] = sum(in[0..i)), out has n + 1 entries; returns total
int64_t dev_exclusive_scan_i32(const int32_t* in, int32_t* out, size_t n, hipStream_t s);
int64_t dev_exclusive_scan_i64(const int64_t* in, int64_t* out, size_t n, hipStream_t s);
void dev_fill_i32(int32_t* p, int32_t v, size_t n, hipStream_t s);
void dev_fill_f64(double* p, double v, size_t n, hipStream_t s);
void dev_fill_f32(float* p, float v, size_t n, hipStream_t s);
void dev_iota_i32(int32_t* p, size_t n, hipStream_t s);
void dev_f64_to_f32(const double* in, float* out, size_t n, hipStream_t s);

// ---- mesh, pattern and assembly (assembly.hip) ----
// reorder + v2e + pattern + values from the caller's host arrays (user
// numbering; kappa may be null = 1); validates them on the device
void build_mesh_device(Handle& h, const double* xyz, const int32_t* conn, const double* kappa);
// bit 0: vertex index out of range, bit 1: NaN coefficient, bit 2: NaN coordinate
int check_mesh_device(int64_t nconn, const int32_t* conn, int64_t nv, int64_t ne, const double* kappa, int64_t nx,
                      const double* xyz, hipStream_t s);
void assemble_values(Handle& h);            // values only, pattern fixed
void element_gradient(Handle& h, const double* a_int, const double* u_int, double* g_user);
void gather_user_to_int(Handle& h, const double* u, double* i, int k);   // i[r] = u[perm[r]]
// z = M r (user numbering, device pointers) through the static device scratch
// in solve.hip: no device allocation; needs a built hierarchy and work vectors
void apply_preconditioner_noalloc(Handle& h, const double* r_user, double* z_user);
int64_t noalloc_max_rows();
// resolves the scratch address of the handle's device when it is current
void noalloc_init(const Handle& h);
void scatter_int_to_user(Handle& h, const double* i, double* u, int k);  // u[perm[r]] = i[r]

// ---- sparse building blocks (sparse.hip) ----
// C = L * Rm with optional column map on Rm's columns and optional removal of
// the diagonal; pattern only (sorted unique columns).
void spgemm_symbolic(const Csr& L, const int32_t* r_rp, const int32_t* r_ci, int64_t r_rows,
                     const int32_t* colmap, int64_t ncols, bool drop_diag, Csr& C, hipStream_t s);
// values of C = L * Rm (FP64) for the pattern found by spgemm_symbolic
void spgemm_numeric(const Csr& L, const Csr& Rm, Csr& C, hipStream_t s);
void csr_transpose(const Csr& P, Csr& R, hipStream_t s);
// (re)build the sliced-ELL copy of A from its CSR arrays (FP64, and FP32 when
// with_f32) with `lanes` (1, 2, 4, 8, 16 or 32) lanes per row; split_diag
// (square A, one lane per row, row order) keeps the diagonal outside the slices
void build_sell(Csr& A, bool with_f32, hipStream_t s, int64_t sort_window = 0, bool with_f16 = false,
                bool with_f64 = true, int lanes = 1, bool split_diag = false);
// LDS staging lists for A's sliced copy (blocks of bs slices); leaves A
// unstaged (lds_bs = 0) when a block holds more than 16384 stored entries or
// more than 8192 distinct columns
void build_lds(Csr& A, int bs, hipStream_t s);
void spmv64(const Csr& A, const double* x, double* y, hipStream_t s);

// ---- AMG setup (amg_setup.hip) ----
void build_hierarchy(Handle& h);
// AggMIS merge-split conditioning of an aggregation to sizes [min_size,
// max_size] (host; aggmis.cpp); rewrites agg, returns the aggregate count
int64_t condition_aggregates(int64_t n, const int32_t* rp, const int32_t* ci, std::vector<int32_t>& agg,
                             int64_t nagg, int min_size, int max_size);
// aggregatorType 4 / 5 (host; aggmis.cpp): flood-fill distance-k MIS from a
// start node derived from the seed, aggregation to the nearest root
int64_t flood_fill_starter(int64_t n, unsigned long long seed);
std::vector<int32_t> flood_fill_mis(int64_t n, const int32_t* rp, const int32_t* ci, int k, int64_t starter);
int64_t aggregate_to_nearest(int64_t n, const int32_t* rp, const int32_t* ci, const std::vector<int32_t>& roots,
                             std::vector<int32_t>& agg);
void aggregate_graph(int64_t n, const int32_t* adj_ptr, const int32_t* adj, int depth, int min_size,
                     int rule, unsigned long long seed, int32_t* agg_out, int64_t* n_agg,
                     hipStream_t s);

// ---- solve (solve.hip) ----
void ensure_krylov(Handle& h, int k);
int pcg_solve(Handle& h, int k, fem_solve_info* info);   // uses h.kb (rhs), h.kx (x0 / out), internal order
void apply_vcycle_fp64(Handle& h, const double* r, double* z);  // level-0 preconditioner, 1 column
void apply_operator_int(Handle& h, const double* x, double* y, int k);
// level lvl's operator (which 0), prolongator (1) or restriction (2)
const Csr& level_matrix_of(const Handle& h, int lvl, int which);
// y = M x for that matrix through the one-column solve kernels, in the
// precision the V-cycle applies it (device FP64 arrays in and out)
// (which 3: the coarsest level's direct solve, lvl must be the coarsest)
void apply_level_int(Handle& h, int lvl, int which, const double* x, double* y);
// the smoother data of a level: rho(D^-1 A) estimate, Chebyshev interval, Jacobi weight
void level_smoother_info(const Handle& h, int lvl, double* rho, double* lmin, double* lmax, double* omega);
// one smoothing step on level lvl through the V-cycle's kernel (device FP64 arrays)
void smoother_step_int(Handle& h, int lvl, int mode, double c1, double c2, const double* x, const double* b,
                       double* d, double* y);

// Profiling: when enabled, HIP events bracket each launch of a kernel family
// on the handle's stream; they are read back (without stalling the stream
// between launches) by flush_profile.
struct ProfScope {
  Handle& h;
  const char* fam;
  double bytes;
  bool on;
  hipEvent_t a = nullptr;
  ProfScope(Handle& hh, const char* f, double b);
  ~ProfScope();
};
void flush_profile(Handle& h);
void release_profile_events(Handle& h);

}  // namespace fem
