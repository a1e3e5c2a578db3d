// C ABI of libfem_mi355x: argument checking, handle lifetime, host <-> device
// staging and the user <-> internal renumbering at the boundary.  Every entry
// point converts exceptions into status codes; nothing calls exit().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "fem_internal.hpp"
#include "mesh_io.hpp"

namespace fem {
void set_kappa_user(Handle& h, const double* kappa_user_dev);
}

using namespace fem;

namespace {

thread_local std::string g_create_error;

template <typename F>
int guarded(fem_handle* hh, F&& f) {
  Handle* h = reinterpret_cast<Handle*>(hh);
  try {
    return f();
  } catch (const FemError& e) {
    if (h) h->err = e.what(); else g_create_error = e.what();
    return e.status;
  } catch (const std::bad_alloc&) {
    if (h) h->err = "host out of memory"; else g_create_error = "host out of memory";
    return FEM_ERR_CAPACITY;
  } catch (const std::exception& e) {
    if (h) h->err = e.what(); else g_create_error = e.what();
    return FEM_ERR_ARG;
  }
}

// Host readers called outside guarded(): an exception (a failed host
// allocation, a stream error) becomes the reader's own failure code.
template <typename F>
int reader_guarded(int fail_code, std::string& err, F&& f) {
  try {
    return f();
  } catch (const std::bad_alloc&) {
    err = "host out of memory while reading the file";
  } catch (const std::exception& e) {
    err = std::string("error while reading the file: ") + e.what();
  }
  return fail_code;
}

Handle* H(fem_handle* h) { return reinterpret_cast<Handle*>(h); }
const Handle* H(const fem_handle* h) { return reinterpret_cast<const Handle*>(h); }

void init_device(Handle& h) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) fail(FEM_ERR_HIP, "no HIP device available");
  if (h.prm.device < 0 || h.prm.device >= count) fail(FEM_ERR_ARG, "device index out of range");
  h.device = h.prm.device;
  HIP_CHECK(hipSetDevice(h.device));
  HIP_CHECK(hipStreamCreateWithFlags(&h.stream, hipStreamNonBlocking));
  h.own_stream = true;
}

// material label -> element coefficient (perform_element_loop_3D.cuh:585-610)
double label_coefficient(int label) {
  if (label >= 2 && label <= 6) return (double)label;
  return 1.0;
}

void create_common(Handle& h, int dim, int64_t nv, const double* xyz, int64_t ne, const int32_t* conn,
                   const double* kappa, double mass, int flags) {
  if (dim != 2 && dim != 3) fail(FEM_ERR_ARG, "dim must be 2 or 3");
  if (nv <= 0 || ne <= 0 || !xyz || !conn) fail(FEM_ERR_ARG, "empty mesh");
  h.dim = dim;
  h.nv = nv;
  h.ne = ne;
  h.mass_coeff = mass;
  h.flags = flags;
  init_device(h);
  {
    DevPoolScope pool;  // the mesh build's temporaries reuse each other's blocks
    build_mesh_device(h, xyz, conn, kappa);  // validates indices and NaNs on the device
    HIP_CHECK(hipStreamSynchronize(h.stream));
  }
}

// internal CSR -> user-numbered CSR (host)
void export_user_csr(const Handle& h, std::vector<int64_t>& rp, std::vector<int32_t>& ci, std::vector<double>& v) {
  const int64_t n = h.A.n, nnz = h.A.nnz;
  std::vector<int32_t> irp(n + 1), ici(nnz), perm(n);
  std::vector<double> iv(nnz);
  h.A.rp.download(irp.data(), n + 1, h.stream);
  h.A.ci.download(ici.data(), nnz, h.stream);
  h.A.v64.download(iv.data(), nnz, h.stream);
  h.perm.download(perm.data(), n, h.stream);
  HIP_CHECK(hipStreamSynchronize(h.stream));
  std::vector<int32_t> iperm(n);
  for (int64_t r = 0; r < n; ++r) iperm[perm[r]] = (int32_t)r;
  rp.assign(n + 1, 0);
  for (int64_t u = 0; u < n; ++u) {
    int32_t r = iperm[u];
    rp[u + 1] = rp[u] + (irp[r + 1] - irp[r]);
  }
  ci.assign(nnz, 0);
  v.assign(nnz, 0.0);
  std::vector<std::pair<int32_t, double>> row;
  for (int64_t u = 0; u < n; ++u) {
    int32_t r = iperm[u];
    row.clear();
    for (int q = irp[r]; q < irp[r + 1]; ++q) row.emplace_back(perm[ici[q]], iv[q]);
    std::sort(row.begin(), row.end(), [](const std::pair<int32_t, double>& a, const std::pair<int32_t, double>& b) {
      return a.first < b.first;
    });
    for (size_t t = 0; t < row.size(); ++t) {
      ci[rp[u] + t] = row[t].first;
      v[rp[u] + t] = row[t].second;
    }
  }
}

// user-numbered CSR (host) -> internal device CSR in h.A
void import_user_csr(Handle& h, const std::vector<int64_t>& rp, const std::vector<int32_t>& ci,
                     const std::vector<double>& v) {
  const int64_t n = (int64_t)rp.size() - 1;
  std::vector<int32_t> perm(n);
  h.perm.download(perm.data(), n, h.stream);
  HIP_CHECK(hipStreamSynchronize(h.stream));
  std::vector<int32_t> iperm(n);
  for (int64_t r = 0; r < n; ++r) iperm[perm[r]] = (int32_t)r;
  const int64_t nnz = rp[n];
  if (nnz >= ((int64_t)1 << 31)) fail(FEM_ERR_CAPACITY, "matrix has 2^31 or more nonzeros");
  std::vector<int32_t> irp(n + 1, 0), ici(nnz);
  std::vector<double> iv(nnz);
  for (int64_t r = 0; r < n; ++r) irp[r + 1] = irp[r] + (int32_t)(rp[perm[r] + 1] - rp[perm[r]]);
  std::vector<std::pair<int32_t, double>> row;
  for (int64_t r = 0; r < n; ++r) {
    int64_t u = perm[r];
    row.clear();
    for (int64_t q = rp[u]; q < rp[u + 1]; ++q) row.emplace_back(iperm[ci[q]], v[q]);
    std::sort(row.begin(), row.end(), [](const std::pair<int32_t, double>& a, const std::pair<int32_t, double>& b) {
      return a.first < b.first;
    });
    for (size_t t = 0; t < row.size(); ++t) {
      ici[irp[r] + t] = row[t].first;
      iv[irp[r] + t] = row[t].second;
    }
  }
  h.A.n = h.A.m = n;
  h.A.nnz = nnz;
  h.A.rp.upload(irp.data(), n + 1, h.stream);
  h.A.ci.upload(ici.data(), nnz, h.stream);
  h.A.v64.upload(iv.data(), nnz, h.stream);
  h.A.v32.release();
  h.A.nslices = 0;
  HIP_CHECK(hipStreamSynchronize(h.stream));
}

bool params_shape_changed(const fem_params& a, const fem_params& b) {
  return a.maxLevels != b.maxLevels || a.topSize != b.topSize || a.randMisParameters != b.randMisParameters ||
         a.proOmega != b.proOmega || a.precision != b.precision || a.aggregationRule != b.aggregationRule ||
         a.seed != b.seed || a.aggregatorType != b.aggregatorType || a.strengthThreshold != b.strengthThreshold;
}

void ensure_setup(Handle& h) {
  if (h.A.n == 0) fail(FEM_ERR_EMPTY, "Error no matrix specified");
  if (!h.H.built) {
    DevPoolScope pool;  // the setup's temporaries reuse each other's blocks
    build_hierarchy(h);
  }
}

// stage nv x k user-order data into internal order on the device
void to_internal(Handle& h, const double* src, int on_device, double* dst, int k) {
  const int64_t m = h.A.n * k;
  const double* dsrc = src;
  if (!on_device) {
    HIP_CHECK(hipMemcpyAsync(h.ubuf.get(), src, sizeof(double) * m, hipMemcpyHostToDevice, h.stream));
    dsrc = h.ubuf.get();
  }
  gather_user_to_int(h, dsrc, dst, k);
}

void from_internal(Handle& h, const double* src, double* dst, int on_device, int k) {
  const int64_t m = h.A.n * k;
  if (on_device) {
    scatter_int_to_user(h, src, dst, k);
  } else {
    scatter_int_to_user(h, src, h.ubuf2.get(), k);
    HIP_CHECK(hipMemcpyAsync(dst, h.ubuf2.get(), sizeof(double) * m, hipMemcpyDeviceToHost, h.stream));
  }
  HIP_CHECK(hipStreamSynchronize(h.stream));
}

void fill_info(const Handle& h, fem_solve_info* info, double setup_ms) {
  if (!info) return;
  info->setup_ms = setup_ms;
  info->num_levels = (int)h.H.levels.size();
  info->grid_complexity = h.H.grid_complexity;
  info->operator_complexity = h.H.operator_complexity;
}

}  // namespace

extern "C" {

void fem_params_default(fem_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->verbose = 0;
  p->maxLevels = 100;
  p->maxIters = 100;
  p->preInnerIters = 5;
  p->postInnerIters = 5;
  p->postRelaxes = 1;
  p->cycleIters = 1;
  p->dsType = 0;
  p->topSize = 256;
  p->randMisParameters = 90102;
  p->partitionMaxSize = 512;
  p->aggregatorType = 0;
  p->convergeType = 0;
  p->tolerance = 1e-6;
  p->cycleType = 0;
  p->solverType = 0;
  p->smootherWeight = 1.0;
  p->proOmega = 0.67;
  p->device = 0;
  p->blockSize = 256;
  p->precision = 0;
  p->smootherType = 0;
  p->aggregationRule = 0;
  p->coarseSolve = 0;
  p->seed = 0x5EEDull;
  p->strengthThreshold = 0.0;
  p->launchMode = 1;
}

int fem_create_from_arrays(int dim, int64_t nv, const double* xyz, int64_t ne, const int32_t* conn,
                           const double* kappa, double mass_coeff, int flags, int verbose, fem_handle** out) {
  if (!out) return FEM_ERR_ARG;
  *out = nullptr;
  Handle* h = new Handle();
  fem_params_default(&h->prm);
  h->prm.verbose = verbose;
  int st = guarded(nullptr, [&]() {
    create_common(*h, dim, nv, xyz, ne, conn, kappa, mass_coeff, flags);
    return (int)FEM_OK;
  });
  if (st != FEM_OK) {
    fem_destroy(reinterpret_cast<fem_handle*>(h));
    return st;
  }
  *out = reinterpret_cast<fem_handle*>(h);
  return FEM_OK;
}

int fem_create_from_files(const char* base, int is_tet, int verbose, fem_handle** out) {
  if (!out || !base) return FEM_ERR_ARG;
  *out = nullptr;
  MeshData m;
  std::string err;
  int rc = reader_guarded(2, err, [&]() { return is_tet ? read_tetgen(base, m, err) : read_trimesh(base, m, err); });
  if (rc) {
    g_create_error = err;
    return FEM_ERR_IO;
  }
  const int npe = m.dim + 1;
  int64_t ne = (int64_t)m.conn.size() / npe;
  std::vector<double> kappa;
  rc = reader_guarded(2, err, [&]() {
    kappa.assign(ne, 1.0);
    for (int64_t e = 0; e < ne; ++e) kappa[e] = label_coefficient(m.labels[e]);
    return 0;
  });
  if (rc) {
    g_create_error = err;
    return FEM_ERR_CAPACITY;
  }
  return fem_create_from_arrays(m.dim, (int64_t)m.xyz.size() / 3, m.xyz.data(), ne, m.conn.data(), kappa.data(), 1.0,
                                m.dim == 2 ? 1 : 0, verbose, out);
}

void fem_release_cached_memory(void) { pool_flush_reserve(); }

void fem_destroy(fem_handle* hh) {
  Handle* h = H(hh);
  if (!h) return;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->h_scal) (void)hipHostFree(h->h_scal);
  if (h->h_act) (void)hipHostFree(h->h_act);
  if (h->iter_graph) (void)hipGraphExecDestroy(h->iter_graph);
  release_profile_events(*h);
  hipStream_t s = h->own_stream ? h->stream : nullptr;
  delete h;  // device buffers free themselves
  if (s) (void)hipStreamDestroy(s);
}

const char* fem_last_error(const fem_handle* h) {
  if (!h) return g_create_error.c_str();
  return H(h)->err.c_str();
}

int fem_set_params(fem_handle* hh, const fem_params* p) {
  if (!hh || !p) return FEM_ERR_ARG;
  return guarded(hh, [&]() {
    Handle& h = *H(hh);
    if (p->maxIters < 0 || p->topSize < 1 || p->maxLevels < 1 || p->preInnerIters < 0 || p->postInnerIters < 0 ||
        !(p->tolerance >= 0) || p->precision < 0 || p->precision > 2 || p->smootherType < 0 ||
        p->smootherType > 1 || p->aggregationRule < 0 || p->aggregationRule > 1 || !(p->strengthThreshold >= 0) ||
        p->strengthThreshold >= 1.0 || p->launchMode < 0 || p->launchMode > 1 || p->cycleType < 0 ||
        p->cycleType > 3)
      fail(FEM_ERR_ARG, "invalid solver parameter");
    if (p->aggregatorType != 0 && (p->aggregatorType < 3 || p->aggregatorType > 5))
      fail(FEM_ERR_ARG, "aggregatorType: 0 (randomized MIS), 3 (AggMIS merge-split), 4 and 5 (flood-fill MIS, "
                        "with and without merge-split) are implemented; 1 and 2 need METIS");
    if (p->aggregatorType >= 3) {
      const int lo = (p->randMisParameters / 1000) % 1000, hi = p->randMisParameters % 1000;
      if (p->randMisParameters < 0 || lo < 1 || hi < lo)
        fail(FEM_ERR_ARG, "aggregatorType 3-5 read aggregate sizes [p / 1000 % 1000, p % 1000] from "
                          "randMisParameters; they must satisfy 1 <= min <= max");
    }
    if (params_shape_changed(h.prm, *p)) h.H.built = false;
    if (p->device != h.prm.device) fail(FEM_ERR_ARG, "device cannot change after construction");
    h.prm = *p;
    ++h.gen;  // smoother coefficients and sweep counts are baked into a captured iteration
    return (int)FEM_OK;
  });
}

int fem_get_params(const fem_handle* h, fem_params* p) {
  if (!h || !p) return FEM_ERR_ARG;
  *p = H(h)->prm;
  return FEM_OK;
}

int fem_set_stream(fem_handle* hh, void* stream) {
  if (!hh) return FEM_ERR_ARG;
  return guarded(hh, [&]() {
    Handle& h = *H(hh);
    HIP_CHECK(hipStreamSynchronize(h.stream));
    if (stream) {
      if (h.own_stream) HIP_CHECK(hipStreamDestroy(h.stream));
      h.stream = (hipStream_t)stream;
      h.own_stream = false;
    } else if (!h.own_stream) {
      HIP_CHECK(hipStreamCreateWithFlags(&h.stream, hipStreamNonBlocking));
      h.own_stream = true;
    }
    return (int)FEM_OK;
  });
}

int64_t fem_get_matrix_rows(const fem_handle* h) { return h ? H(h)->A.n : 0; }
int64_t fem_get_matrix_nnz(const fem_handle* h) { return h ? H(h)->A.nnz : 0; }
int64_t fem_get_num_elements(const fem_handle* h) { return h ? H(h)->ne : 0; }

int fem_get_matrix_csr(const fem_handle* hc, int64_t* row_ptr, int32_t* cols, double* vals) {
  if (!hc) return FEM_ERR_ARG;
  fem_handle* hh = const_cast<fem_handle*>(hc);
  return guarded(hh, [&]() {
    std::vector<int64_t> rp;
    std::vector<int32_t> ci;
    std::vector<double> v;
    export_user_csr(*H(hh), rp, ci, v);
    if (row_ptr) std::copy(rp.begin(), rp.end(), row_ptr);
    if (cols) std::copy(ci.begin(), ci.end(), cols);
    if (vals) std::copy(v.begin(), v.end(), vals);
    return (int)FEM_OK;
  });
}

int fem_read_matlab_sparse(fem_handle* hh, const char* filename) {
  if (!hh || !filename) return FEM_ERR_ARG;
  Handle& h = *H(hh);
  SparseCoo A;
  std::string err;
  int rc = reader_guarded(1, err, [&]() { return read_matlab_sparse(filename, A, err); });
  if (rc != 0) {
    h.err = err;
    return rc;
  }
  return guarded(hh, [&]() {
    const int64_t n = h.A.n;
    if (A.nrows != n || A.ncols != n)
      fail(FEM_ERR_ARG, "matrix file is " + std::to_string(A.nrows) + " x " + std::to_string(A.ncols) +
                            " but the mesh has " + std::to_string(n) + " vertices");
    const bool round_float = (h.flags & 2) != 0;
    // mesh pattern (1e-12 on every slot) + file entries, summed (cusp::add)
    std::vector<int64_t> mrp;
    std::vector<int32_t> mci;
    std::vector<double> mv;
    export_user_csr(h, mrp, mci, mv);
    std::vector<std::vector<std::pair<int32_t, double>>> rows(n);
    const double tiny = round_float ? (double)1e-12f : 1e-12;
    for (int64_t u = 0; u < n; ++u)
      for (int64_t q = mrp[u]; q < mrp[u + 1]; ++q) rows[u].emplace_back(mci[q], tiny);
    for (size_t t = 0; t < A.vals.size(); ++t) {
      double a = round_float ? (double)(float)A.vals[t] : A.vals[t];
      rows[A.rows[t]].emplace_back(A.cols[t], a);
    }
    std::vector<int64_t> rp(n + 1, 0);
    std::vector<int32_t> ci;
    std::vector<double> v;
    for (int64_t u = 0; u < n; ++u) {
      auto& r = rows[u];
      std::stable_sort(r.begin(), r.end(), [](const std::pair<int32_t, double>& a,
                                              const std::pair<int32_t, double>& b) { return a.first < b.first; });
      for (size_t t = 0; t < r.size();) {
        size_t e = t;
        double sum = 0.0;
        while (e < r.size() && r[e].first == r[t].first) sum += r[e++].second;
        if (round_float) sum = (double)(float)sum;
        ci.push_back(r[t].first);
        v.push_back(sum);
        t = e;
      }
      rp[u + 1] = (int64_t)ci.size();
    }
    import_user_csr(h, rp, ci, v);
    h.matrix_from_file = true;
    h.H.built = false;
    return (int)FEM_OK;
  });
}

int fem_update_kappa(fem_handle* hh, const double* kappa) {
  if (!hh || !kappa) return FEM_ERR_ARG;
  return guarded(hh, [&]() {
    Handle& h = *H(hh);
    if (!h.has_mesh || h.matrix_from_file) fail(FEM_ERR_STATE, "operator is not assembled from a mesh");
    DevBuf<double> ku(h.ne);
    const HostPiece pc = {ku.get(), kappa, sizeof(double) * (size_t)h.ne};
    upload_pieces(&pc, 1, h.stream);
    if (check_mesh_device(0, nullptr, 0, h.ne, ku.get(), 0, nullptr, h.stream) & 2)
      fail(FEM_ERR_NUMERIC, "NaN in element coefficient");
    set_kappa_user(h, ku.get());
    assemble_values(h);
    HIP_CHECK(hipStreamSynchronize(h.stream));
    h.H.built = false;
    return (int)FEM_OK;
  });
}

int fem_setup(fem_handle* hh) {
  if (!hh) return FEM_ERR_ARG;
  return guarded(hh, [&]() {
    Handle& h = *H(hh);
    ensure_setup(h);
    ensure_krylov(h, 1);  // one-column work vectors (what fem_apply_preconditioner_noalloc needs)
    noalloc_init(h);
    return (int)FEM_OK;
  });
}

int fem_solve(fem_handle* hh, const double* b, double* x, int nrhs, int on_device, fem_solve_info* info) {
  if (!hh || !b || !x || nrhs < 1 || nrhs > 64) return FEM_ERR_ARG;
  return guarded(hh, [&]() {
    Handle& h = *H(hh);
    if (info) std::memset(info, 0, sizeof(*info));
    double setup_ms = 0.0;
    if (h.A.n == 0) fail(FEM_ERR_EMPTY, "Error no matrix specified");
    if (!h.H.built) {
      ensure_setup(h);
      setup_ms = h.H.setup_ms;
    }
    ensure_krylov(h, nrhs);
    to_internal(h, b, on_device, h.kb.get(), nrhs);
    to_internal(h, x, on_device, h.kx.get(), nrhs);
    int st = pcg_solve(h, nrhs, info);
    from_internal(h, h.kx.get(), x, on_device, nrhs);
    fill_info(h, info, setup_ms);
    if (st == FEM_ERR_NOT_CONVERGED) h.err = "solve reached maxIters above tolerance";
    return st;
  });
}

int fem_adjoint_grad(fem_handle* hh, const double* u, const double* dLdu, double* dLdkappa, double* lambda_out,
                     int on_device, fem_solve_info* info) {
  if (!hh || !u || !dLdu || !dLdkappa) return FEM_ERR_ARG;
  return guarded(hh, [&]() {
    Handle& h = *H(hh);
    if (!h.has_mesh || h.matrix_from_file) fail(FEM_ERR_STATE, "adjoint needs an operator assembled from a mesh");
    if (info) std::memset(info, 0, sizeof(*info));
    double setup_ms = 0.0;
    if (!h.H.built) {
      ensure_setup(h);
      setup_ms = h.H.setup_ms;
    }
    ensure_krylov(h, 1);
    // K lambda = dL/du (K symmetric), zero initial guess
    to_internal(h, dLdu, on_device, h.kb.get(), 1);
    HIP_CHECK(hipMemsetAsync(h.kx.get(), 0, sizeof(double) * h.A.n, h.stream));
    int st = pcg_solve(h, 1, info);
    // u into internal order (kz is free after the solve)
    to_internal(h, u, on_device, h.kz.get(), 1);
    double* g = dLdkappa;
    DevBuf<double> gdev;
    if (!on_device) {
      gdev.resize(h.ne);
      g = gdev.get();
    }
    element_gradient(h, h.kx.get(), h.kz.get(), g);
    if (!on_device) {
      HIP_CHECK(hipMemcpyAsync(dLdkappa, g, sizeof(double) * h.ne, hipMemcpyDeviceToHost, h.stream));
    }
    if (lambda_out) from_internal(h, h.kx.get(), lambda_out, on_device, 1);
    HIP_CHECK(hipStreamSynchronize(h.stream));
    fill_info(h, info, setup_ms);
    return st;
  });
}

int fem_element_grad(fem_handle* hh, const double* a, const double* u, double* dLdkappa, int on_device) {
  if (!hh || !a || !u || !dLdkappa) return FEM_ERR_ARG;
  return guarded(hh, [&]() {
    Handle& h = *H(hh);
    if (!h.has_mesh) fail(FEM_ERR_STATE, "no mesh");
    ensure_krylov(h, 1);
    to_internal(h, a, on_device, h.kq.get(), 1);
    to_internal(h, u, on_device, h.kz.get(), 1);
    double* g = dLdkappa;
    DevBuf<double> gdev;
    if (!on_device) {
      gdev.resize(h.ne);
      g = gdev.get();
    }
    element_gradient(h, h.kq.get(), h.kz.get(), g);
    if (!on_device) HIP_CHECK(hipMemcpyAsync(dLdkappa, g, sizeof(double) * h.ne, hipMemcpyDeviceToHost, h.stream));
    HIP_CHECK(hipStreamSynchronize(h.stream));
    return (int)FEM_OK;
  });
}

int fem_apply_operator(fem_handle* hh, const double* x, double* y, int nrhs, int on_device) {
  if (!hh || !x || !y || nrhs < 1 || nrhs > 64) return FEM_ERR_ARG;
  return guarded(hh, [&]() {
    Handle& h = *H(hh);
    if (h.A.n == 0) fail(FEM_ERR_EMPTY, "Error no matrix specified");
    ensure_krylov(h, nrhs);
    to_internal(h, x, on_device, h.kp.get(), nrhs);
    apply_operator_int(h, h.kp.get(), h.kq.get(), nrhs);
    from_internal(h, h.kq.get(), y, on_device, nrhs);
    return (int)FEM_OK;
  });
}

int fem_apply_preconditioner(fem_handle* hh, const double* r, double* z, int on_device) {
  if (!hh || !r || !z) return FEM_ERR_ARG;
  return guarded(hh, [&]() {
    Handle& h = *H(hh);
    ensure_setup(h);
    ensure_krylov(h, 1);
    to_internal(h, r, on_device, h.kr.get(), 1);
    apply_vcycle_fp64(h, h.kr.get(), h.kz.get());
    from_internal(h, h.kz.get(), z, on_device, 1);
    return (int)FEM_OK;
  });
}

int fem_apply_preconditioner_noalloc(fem_handle* hh, const double* r, double* z) {
  if (!hh || !r || !z) return FEM_ERR_ARG;
  return guarded(hh, [&]() {
    apply_preconditioner_noalloc(*H(hh), r, z);
    return (int)FEM_OK;
  });
}

int64_t fem_noalloc_max_rows(void) { return noalloc_max_rows(); }

int fem_apply_level(fem_handle* hh, int lvl, int which, const double* x, double* y) {
  if (!hh || !x || !y) return FEM_ERR_ARG;
  return guarded(hh, [&]() {
    Handle& h = *H(hh);
    ensure_setup(h);
    const Csr& M = level_matrix_of(h, lvl, which);
    DevBuf<double> dx(M.m > 0 ? M.m : 1), dy(M.n > 0 ? M.n : 1);
    dx.upload(x, M.m, h.stream);
    if (which < 0 || which > 3) fail(FEM_ERR_ARG, "which: 0 operator, 1 prolongator, 2 restriction, 3 coarse solve");
    apply_level_int(h, lvl, which, dx.get(), dy.get());
    dy.download(y, M.n, h.stream);
    HIP_CHECK(hipStreamSynchronize(h.stream));
    return (int)FEM_OK;
  });
}

int fem_get_level_smoother(const fem_handle* hc, int lvl, double* rho, double* lmin, double* lmax, double* omega) {
  if (!hc) return FEM_ERR_ARG;
  fem_handle* hh = const_cast<fem_handle*>(hc);
  return guarded(hh, [&]() {
    level_smoother_info(*H(hh), lvl, rho, lmin, lmax, omega);
    return (int)FEM_OK;
  });
}

int fem_apply_smoother_step(fem_handle* hh, int lvl, int mode, double c1, double c2, const double* x,
                            const double* b, double* d, double* y) {
  if (!hh || !x || !b || !d || !y) return FEM_ERR_ARG;
  return guarded(hh, [&]() {
    Handle& h = *H(hh);
    ensure_setup(h);
    const Csr& A = level_matrix_of(h, lvl, 0);
    const int64_t n = A.n;
    DevBuf<double> dx(n), db(n), dd(n), dy(n);
    dx.upload(x, n, h.stream);
    db.upload(b, n, h.stream);
    dd.upload(d, n, h.stream);
    smoother_step_int(h, lvl, mode, c1, c2, dx.get(), db.get(), dd.get(), dy.get());
    dy.download(y, n, h.stream);
    dd.download(d, n, h.stream);
    HIP_CHECK(hipStreamSynchronize(h.stream));
    return (int)FEM_OK;
  });
}

int fem_get_num_levels(const fem_handle* h) { return h ? (int)H(h)->H.levels.size() : 0; }

int fem_get_level_info(const fem_handle* hc, int lvl, int64_t* n, int64_t* nnz) {
  if (!hc) return FEM_ERR_ARG;
  const Handle& h = *H(hc);
  if (lvl < 0 || lvl >= (int)h.H.levels.size()) return FEM_ERR_ARG;
  const Csr& A = lvl == 0 ? h.A : h.H.levels[lvl]->A;
  if (n) *n = A.n;
  if (nnz) *nnz = A.nnz;
  return FEM_OK;
}

int fem_get_level_matrix(const fem_handle* hc, int lvl, int32_t* row_ptr, int32_t* cols, double* vals,
                         int32_t* aggregates) {
  if (!hc) return FEM_ERR_ARG;
  fem_handle* hh = const_cast<fem_handle*>(hc);
  return guarded(hh, [&]() {
    Handle& h = *H(hh);
    if (lvl < 0 || lvl >= (int)h.H.levels.size()) fail(FEM_ERR_ARG, "level out of range");
    const Csr& A = lvl == 0 ? h.A : h.H.levels[lvl]->A;
    if (row_ptr) A.rp.download(row_ptr, A.n + 1, h.stream);
    if (cols) A.ci.download(cols, A.nnz, h.stream);
    if (vals) A.v64.download(vals, A.nnz, h.stream);
    if (aggregates && h.H.levels[lvl]->agg.size()) h.H.levels[lvl]->agg.download(aggregates, A.n, h.stream);
    HIP_CHECK(hipStreamSynchronize(h.stream));
    return (int)FEM_OK;
  });
}

int fem_get_level_prolongator(const fem_handle* hc, int lvl, int64_t* nnz, int32_t* row_ptr, int32_t* cols,
                              double* vals) {
  if (!hc) return FEM_ERR_ARG;
  fem_handle* hh = const_cast<fem_handle*>(hc);
  return guarded(hh, [&]() {
    Handle& h = *H(hh);
    if (lvl < 0 || lvl >= (int)h.H.levels.size() || lvl == h.H.coarse) fail(FEM_ERR_ARG, "level has no prolongator");
    const Csr& P = h.H.levels[lvl]->P;
    if (nnz) *nnz = P.nnz;
    if (row_ptr) P.rp.download(row_ptr, P.n + 1, h.stream);
    if (cols) P.ci.download(cols, P.nnz, h.stream);
    if (vals) P.v64.download(vals, P.nnz, h.stream);
    HIP_CHECK(hipStreamSynchronize(h.stream));
    return (int)FEM_OK;
  });
}

int fem_get_permutation(const fem_handle* hc, int32_t* perm) {
  if (!hc || !perm) return FEM_ERR_ARG;
  fem_handle* hh = const_cast<fem_handle*>(hc);
  return guarded(hh, [&]() {
    Handle& h = *H(hh);
    h.perm.download(perm, h.A.n, h.stream);
    HIP_CHECK(hipStreamSynchronize(h.stream));
    return (int)FEM_OK;
  });
}

int fem_device_aggregate(int64_t n, const int32_t* adj_ptr, const int32_t* adj, int depth, int min_size, int rule,
                         unsigned long long seed, int32_t* aggregate, int32_t* n_agg) {
  if (n < 0 || !adj_ptr || !aggregate || !n_agg) return FEM_ERR_ARG;
  return guarded(nullptr, [&]() {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) fail(FEM_ERR_HIP, "no HIP device available");
    hipStream_t s;
    HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    int64_t m = adj_ptr[n];
    for (int64_t t = 0; t < m; ++t)
      if (adj[t] < 0 || adj[t] >= n) fail(FEM_ERR_ARG, "adjacency index out of range");
    DevBuf<int32_t> p, a, g;
    p.upload(adj_ptr, n + 1, s);
    a.upload(adj, m > 0 ? m : 0, s);
    g.resize(n > 0 ? n : 1);
    int64_t na = 0;
    aggregate_graph(n, p.get(), a.get(), depth, min_size, rule, seed, g.get(), &na, s);
    g.download(aggregate, n, s);
    HIP_CHECK(hipStreamSynchronize(s));
    HIP_CHECK(hipStreamDestroy(s));
    *n_agg = (int32_t)na;
    return (int)FEM_OK;
  });
}

int fem_condition_aggregates(int64_t n, const int32_t* adj_ptr, const int32_t* adj, int32_t* aggregate,
                             int64_t n_agg, int min_size, int max_size, int64_t* n_agg_out) {
  if (n < 0 || !adj_ptr || !aggregate || !n_agg_out || n_agg < 0 || min_size < 1 || max_size < min_size)
    return FEM_ERR_ARG;
  return guarded(nullptr, [&]() {
    const int64_t m = adj_ptr[n];
    if (m > 0 && !adj) fail(FEM_ERR_ARG, "adjacency missing");
    for (int64_t t = 0; t < m; ++t)
      if (adj[t] < 0 || adj[t] >= n) fail(FEM_ERR_ARG, "adjacency index out of range");
    std::vector<int32_t> agg(aggregate, aggregate + n);
    std::vector<char> used((size_t)n_agg, 0);
    for (int64_t i = 0; i < n; ++i) {
      if (agg[(size_t)i] < 0 || agg[(size_t)i] >= n_agg) fail(FEM_ERR_ARG, "aggregate id out of range");
      used[(size_t)agg[(size_t)i]] = 1;
    }
    for (int64_t a = 0; a < n_agg; ++a)
      if (!used[(size_t)a]) fail(FEM_ERR_ARG, "aggregate ids must be dense (every id in use)");
    *n_agg_out = condition_aggregates(n, adj_ptr, adj, agg, n_agg, min_size, max_size);
    std::copy(agg.begin(), agg.end(), aggregate);
    return (int)FEM_OK;
  });
}

int fem_flood_fill_aggregate(int64_t n, const int32_t* adj_ptr, const int32_t* adj, int depth,
                             unsigned long long seed, int64_t* starter_out, int32_t* roots_out,
                             int32_t* aggregate_out, int64_t* n_agg_out) {
  if (n < 0 || !adj_ptr || depth < 0 || !roots_out || !aggregate_out || !n_agg_out) return FEM_ERR_ARG;
  return guarded(nullptr, [&]() {
    const int64_t m = adj_ptr[n];
    if (m > 0 && !adj) fail(FEM_ERR_ARG, "adjacency missing");
    for (int64_t i = 0; i < n; ++i)
      if (adj_ptr[i] > adj_ptr[i + 1]) fail(FEM_ERR_ARG, "row offsets must not decrease");
    for (int64_t t = 0; t < m; ++t)
      if (adj[t] < 0 || adj[t] >= n) fail(FEM_ERR_ARG, "adjacency index out of range");
    const int64_t start = flood_fill_starter(n, seed);
    if (starter_out) *starter_out = start;
    const std::vector<int32_t> roots = flood_fill_mis(n, adj_ptr, adj, depth, start);
    std::vector<int32_t> agg;
    *n_agg_out = aggregate_to_nearest(n, adj_ptr, adj, roots, agg);
    std::copy(roots.begin(), roots.end(), roots_out);
    std::copy(agg.begin(), agg.end(), aggregate_out);
    return (int)FEM_OK;
  });
}

int fem_profile_enable(fem_handle* hh, int enable) {
  if (!hh) return FEM_ERR_ARG;
  return guarded(hh, [&]() {
    Handle& h = *H(hh);
    flush_profile(h);
    h.profiling = enable != 0;
    if (enable) h.stats.clear();
    return (int)FEM_OK;
  });
}

int fem_profile_get(const fem_handle* hc, const char* family, int64_t* launches, double* total_ms,
                    double* bytes_per_launch) {
  if (!hc || !family) return FEM_ERR_ARG;
  Handle& h = *H(const_cast<fem_handle*>(hc));
  int st = guarded(const_cast<fem_handle*>(hc), [&]() {
    flush_profile(h);
    return (int)FEM_OK;
  });
  if (st != FEM_OK) return st;
  auto it = h.stats.find(family);
  if (it == h.stats.end()) {
    if (launches) *launches = 0;
    if (total_ms) *total_ms = 0;
    if (bytes_per_launch) *bytes_per_launch = 0;
    return FEM_OK;
  }
  if (launches) *launches = it->second.launches;
  if (total_ms) *total_ms = it->second.total_ms;
  if (bytes_per_launch) *bytes_per_launch = it->second.bytes;
  return FEM_OK;
}

int fem_read_tetgen(const char* base, int64_t* nv, double** xyz, int64_t* ne, int32_t** tets, int32_t** labels) {
  if (!base || !nv || !xyz || !ne || !tets) return FEM_ERR_ARG;
  MeshData m;
  std::string err;
  if (reader_guarded(2, err, [&]() { return read_tetgen(base, m, err); })) {
    g_create_error = err;
    return FEM_ERR_IO;
  }
  double* px = (double*)std::malloc(sizeof(double) * (m.xyz.size() ? m.xyz.size() : 1));
  int32_t* pt = (int32_t*)std::malloc(sizeof(int32_t) * (m.conn.size() ? m.conn.size() : 1));
  int32_t* pl = labels ? (int32_t*)std::malloc(sizeof(int32_t) * (m.labels.size() ? m.labels.size() : 1)) : nullptr;
  if (!px || !pt || (labels && !pl)) {
    std::free(px);
    std::free(pt);
    std::free(pl);
    g_create_error = "host out of memory";
    return FEM_ERR_CAPACITY;
  }
  std::copy(m.xyz.begin(), m.xyz.end(), px);
  std::copy(m.conn.begin(), m.conn.end(), pt);
  if (labels) std::copy(m.labels.begin(), m.labels.end(), pl);
  *nv = (int64_t)m.xyz.size() / 3;
  *ne = (int64_t)m.conn.size() / 4;
  *xyz = px;
  *tets = pt;
  if (labels) *labels = pl;
  return FEM_OK;
}

int fem_read_trimesh(const char* filename, int64_t* nv, double** xyz, int64_t* ne, int32_t** tris) {
  if (!filename || !nv || !xyz || !ne || !tris) return FEM_ERR_ARG;
  MeshData m;
  std::string err;
  if (reader_guarded(2, err, [&]() { return read_trimesh(filename, m, err); })) {
    g_create_error = err;
    return FEM_ERR_IO;
  }
  double* px = (double*)std::malloc(sizeof(double) * (m.xyz.size() ? m.xyz.size() : 1));
  int32_t* pt = (int32_t*)std::malloc(sizeof(int32_t) * (m.conn.size() ? m.conn.size() : 1));
  if (!px || !pt) {
    std::free(px);
    std::free(pt);
    g_create_error = "host out of memory";
    return FEM_ERR_CAPACITY;
  }
  std::copy(m.xyz.begin(), m.xyz.end(), px);
  std::copy(m.conn.begin(), m.conn.end(), pt);
  *nv = (int64_t)m.xyz.size() / 3;
  *ne = (int64_t)m.conn.size() / 3;
  *xyz = px;
  *tris = pt;
  return FEM_OK;
}

void fem_free(void* p) { std::free(p); }

int fem_read_matlab_array(const char* filename, double* out, int64_t* n) {
  if (!filename || !n) return -1;
  std::vector<double> v;
  std::string err;
  int rc = reader_guarded(-1, err, [&]() { return read_matlab_array(filename, v, err); });
  if (rc) {
    g_create_error = err;
    return rc;
  }
  *n = (int64_t)v.size();
  if (out) std::copy(v.begin(), v.end(), out);
  return 0;
}

int fem_write_matlab_array(const char* filename, const double* x, int64_t n) {
  if (!filename || (!x && n > 0)) return FEM_ERR_ARG;
  std::string err;
  int rc = write_matlab_array(filename, x, n, err);
  if (rc) g_create_error = err;
  return rc ? FEM_ERR_IO : FEM_OK;
}

int fem_read_matlab_sparse_coo(const char* filename, int round_to_float, int64_t* nrows, int64_t* ncols,
                               int64_t* nnz, int32_t* rows, int32_t* cols, double* vals) {
  if (!filename || !nrows || !ncols || !nnz) return FEM_ERR_ARG;
  SparseCoo A;
  std::string err;
  int rc = reader_guarded(1, err, [&]() { return read_matlab_sparse(filename, A, err); });
  if (rc) {
    g_create_error = err;
    return rc;
  }
  *nrows = A.nrows;
  *ncols = A.ncols;
  *nnz = (int64_t)A.vals.size();
  if (rows) std::copy(A.rows.begin(), A.rows.end(), rows);
  if (cols) std::copy(A.cols.begin(), A.cols.end(), cols);
  if (vals)
    for (size_t t = 0; t < A.vals.size(); ++t) vals[t] = round_to_float ? (double)(float)A.vals[t] : A.vals[t];
  return 0;
}

int fem_write_vtk(const char* fname, int dim, int64_t nv, const double* xyz, int64_t ne, const int32_t* conn,
                  const double* values, int64_t nvalues) {
  if (!fname || !xyz || !conn || (!values && nvalues > 0) || (dim != 2 && dim != 3)) return FEM_ERR_ARG;
  std::string err;
  int rc = write_vtk(fname, dim, nv, xyz, ne, conn, values, nvalues, err);
  if (rc) g_create_error = err;
  return rc ? FEM_ERR_IO : FEM_OK;
}

const char* fem_version(void) { return "fem_mi355x 0.1 (gfx950)"; }

}  // extern "C"
