/*
 * FEMSolver.h -- header-only C++ drop-in for SCI-Solver_FEM's FEMSolver
 * (src/FEMSolver.h:14-64), implemented on the C ABI of libfem_mi355x.so.
 *
 * Same constructor, public parameter fields (names and defaults of
 * src/FEMSolver.cu:11-33) and methods.  Differences a caller can observe:
 *  - errors never call exit(): construction failures throw
 *    std::runtime_error, solveFEM on an empty matrix throws
 *    std::invalid_argument as the reference does (src/FEMSolver.cu:131-139);
 *  - solveFEM returns the solution of A x = b in the caller's vertex order
 *    (the reference permutes A but not b/x at level 0);
 *  - the assembled matrix lives on the GPU; getMatrixCsr() copies it out in
 *    place of the public A_h_ member.
 * Link: -L<repo>/sci-solver_fem_amd/lib -lfem_mi355x
 */
#ifndef FEM_MI355X_FEMSOLVER_H
#define FEM_MI355X_FEMSOLVER_H

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "fem_mi355x.h"

/* host vector type of the reference (cusp::array1d<double, host_memory>);
 * only size / operator[] / push_back / clear are used by its callers */
typedef std::vector<double> Vector_h_CG;

class FEMSolver {
 public:
  FEMSolver(std::string fname = "../src/test/test_data/simple", bool isTetMesh = true, bool verbose = false)
      : verbose_(verbose), filename_(fname), isTetMesh_(isTetMesh), h_(nullptr) {
    fem_params p;
    fem_params_default(&p);
    load_fields(p);
    verbose_ = verbose;
    if (fem_create_from_files(fname.c_str(), isTetMesh ? 1 : 0, verbose ? 1 : 0, &h_) != FEM_OK)
      throw std::runtime_error(std::string("FEMSolver: ") + fem_last_error(nullptr));
  }
  virtual ~FEMSolver() {
    if (h_) fem_destroy(h_);
  }
  FEMSolver(const FEMSolver&) = delete;
  FEMSolver& operator=(const FEMSolver&) = delete;

  /* src/FEMSolver.cu:58-93: x_h = initial guess in, solution out; b_h unchanged */
  void solveFEM(Vector_h_CG* x_h, Vector_h_CG* b_h) {
    checkMatrixForValidContents();
    push_fields();
    const size_t n = getMatrixRows();
    if (!x_h || !b_h || b_h->size() != n) throw std::invalid_argument("solveFEM: b_h must have getMatrixRows() entries");
    if (x_h->size() != n) x_h->assign(n, 0.0);
    fem_solve_info info;
    int st = fem_solve(h_, b_h->data(), x_h->data(), 1, 0, &info);
    if (st != FEM_OK && st != FEM_ERR_NOT_CONVERGED)
      throw std::runtime_error(std::string("solveFEM: ") + fem_last_error(h_));
    last_info_ = info;
    if (verbose_)
      std::printf("solveFEM: %d iterations, relative residual %.3e, %d levels\n", info.iterations,
                  info.rel_residual, info.num_levels);
  }

  /* src/FEMSolver.cu:141-171: the constructor already assembled A; calling this
   * again re-assembles it from the mesh with the reference coefficients */
  void getMatrixFromMesh() {
    std::vector<double> kappa(fem_get_num_elements(h_), 1.0);
    if (isTetMesh_) {
      int64_t nv, ne;
      double* xyz;
      int32_t *tets, *lab;
      if (fem_read_tetgen(filename_.c_str(), &nv, &xyz, &ne, &tets, &lab) != FEM_OK)
        throw std::runtime_error(std::string("getMatrixFromMesh: ") + fem_last_error(nullptr));
      for (int64_t e = 0; e < ne && e < (int64_t)kappa.size(); ++e)
        kappa[e] = (lab[e] >= 2 && lab[e] <= 6) ? (double)lab[e] : 1.0;
      fem_free(xyz);
      fem_free(tets);
      fem_free(lab);
    }
    if (fem_update_kappa(h_, kappa.data()) != FEM_OK)
      throw std::runtime_error(std::string("getMatrixFromMesh: ") + fem_last_error(h_));
  }

  /* src/FEMSolver.cu:177-356: 0 on success, 1 / -1 on a malformed file */
  int readMatlabSparseMatrix(const std::string& filename) {
    int rc = fem_read_matlab_sparse(h_, filename.c_str());
    if (rc != 0) std::fprintf(stderr, "%s\n", fem_last_error(h_));
    return rc;
  }

  /* src/FEMSolver.cu:358-449 */
  int readMatlabArray(const std::string& filename, Vector_h_CG* rhs) {
    int64_t n = 0;
    int rc = fem_read_matlab_array(filename.c_str(), nullptr, &n);
    if (rc != 0) {
      std::fprintf(stderr, "%s\n", fem_last_error(nullptr));
      return rc;
    }
    rhs->assign((size_t)n, 0.0);
    return fem_read_matlab_array(filename.c_str(), rhs->data(), &n);
  }

  /* src/FEMSolver.cu:451-517 */
  int writeMatlabArray(const std::string& filename, const Vector_h_CG& array) {
    return fem_write_matlab_array(filename.c_str(), array.data(), (int64_t)array.size()) == FEM_OK ? 0 : 1;
  }

  /* src/FEMSolver.cu:131-139 */
  void checkMatrixForValidContents() {
    if (fem_get_matrix_rows(h_) == 0) {
      if (verbose_) std::printf("Error no matrix specified\n");
      throw std::invalid_argument("Error no matrix specified");
    }
  }

  /* src/FEMSolver.cu:519-582: writes <fname>.vtk */
  void writeVTK(std::vector<double> values, std::string fname) {
    int64_t nv = 0, ne = 0;
    double* xyz = nullptr;
    int32_t* conn = nullptr;
    int dim = isTetMesh_ ? 3 : 2;
    int rc;
    if (isTetMesh_) {
      int32_t* lab = nullptr;
      rc = fem_read_tetgen(filename_.c_str(), &nv, &xyz, &ne, &conn, &lab);
      if (rc == FEM_OK) fem_free(lab);
    } else {
      rc = fem_read_trimesh(filename_.c_str(), &nv, &xyz, &ne, &conn);
    }
    if (rc != FEM_OK) throw std::runtime_error(std::string("writeVTK: ") + fem_last_error(nullptr));
    fem_write_vtk(fname.c_str(), dim, nv, xyz, ne, conn, values.data(), (int64_t)values.size());
    fem_free(xyz);
    fem_free(conn);
  }

  /* src/FEMSolver.cu:95-97 */
  size_t getMatrixRows() { return (size_t)fem_get_matrix_rows(h_); }

  /* in place of the public A_h_: assembled matrix, CSR, caller's vertex order */
  void getMatrixCsr(std::vector<int64_t>* row_ptr, std::vector<int32_t>* cols, std::vector<double>* vals) {
    const int64_t n = fem_get_matrix_rows(h_), nnz = fem_get_matrix_nnz(h_);
    row_ptr->resize(n + 1);
    cols->resize(nnz);
    vals->resize(nnz);
    if (fem_get_matrix_csr(h_, row_ptr->data(), cols->data(), vals->data()) != FEM_OK)
      throw std::runtime_error(std::string("getMatrixCsr: ") + fem_last_error(h_));
  }

  const fem_solve_info& lastSolveInfo() const { return last_info_; }
  fem_handle* handle() { return h_; }

  // data members (src/FEMSolver.h:40-61)
  bool verbose_;
  std::string filename_;
  int maxLevels_;
  int maxIters_;
  int preInnerIters_;
  int postInnerIters_;
  int postRelaxes_;
  int cycleIters_;
  int dsType_;
  int topSize_;
  int randMisParameters_;
  int partitionMaxSize_;
  int aggregatorType_;
  int convergeType_;
  double tolerance_;
  int cycleType_;
  int solverType_;
  double smootherWeight_;
  double proOmega_;
  int device_;
  int blockSize_;
  // additions
  int precision_;
  int smootherType_;
  int aggregationRule_;
  unsigned long long seed_;
  double strengthThreshold_;
  int launchMode_;

 private:
  void load_fields(const fem_params& p) {
    maxLevels_ = p.maxLevels;
    maxIters_ = p.maxIters;
    preInnerIters_ = p.preInnerIters;
    postInnerIters_ = p.postInnerIters;
    postRelaxes_ = p.postRelaxes;
    cycleIters_ = p.cycleIters;
    dsType_ = p.dsType;
    topSize_ = p.topSize;
    randMisParameters_ = p.randMisParameters;
    partitionMaxSize_ = p.partitionMaxSize;
    aggregatorType_ = p.aggregatorType;
    convergeType_ = p.convergeType;
    tolerance_ = p.tolerance;
    cycleType_ = p.cycleType;
    solverType_ = p.solverType;
    smootherWeight_ = p.smootherWeight;
    proOmega_ = p.proOmega;
    device_ = p.device;
    blockSize_ = p.blockSize;
    precision_ = p.precision;
    smootherType_ = p.smootherType;
    aggregationRule_ = p.aggregationRule;
    seed_ = p.seed;
    strengthThreshold_ = p.strengthThreshold;
    launchMode_ = p.launchMode;
  }
  void push_fields() {
    fem_params p;
    fem_get_params(h_, &p);
    p.verbose = verbose_ ? 1 : 0;
    p.maxLevels = maxLevels_;
    p.maxIters = maxIters_;
    p.preInnerIters = preInnerIters_;
    p.postInnerIters = postInnerIters_;
    p.postRelaxes = postRelaxes_;
    p.cycleIters = cycleIters_;
    p.dsType = dsType_;
    p.topSize = topSize_;
    p.randMisParameters = randMisParameters_;
    p.partitionMaxSize = partitionMaxSize_;
    p.aggregatorType = aggregatorType_;
    p.convergeType = convergeType_;
    p.tolerance = tolerance_;
    p.cycleType = cycleType_;
    p.solverType = solverType_;
    p.smootherWeight = smootherWeight_;
    p.proOmega = proOmega_;
    p.blockSize = blockSize_;
    p.precision = precision_;
    p.smootherType = smootherType_;
    p.aggregationRule = aggregationRule_;
    p.seed = seed_;
    p.strengthThreshold = strengthThreshold_;
    p.launchMode = launchMode_;
    if (fem_set_params(h_, &p) != FEM_OK) throw std::invalid_argument(std::string("FEMSolver: ") + fem_last_error(h_));
  }

  bool isTetMesh_;
  fem_handle* h_;
  fem_solve_info last_info_{};
};

#endif /* FEM_MI355X_FEMSOLVER_H */
