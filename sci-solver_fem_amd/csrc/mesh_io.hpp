// Host-side mesh and MATLAB file IO (no GPU needed).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace fem {

struct MeshData {
  int dim = 3;                  // 3: tets (4 per element), 2: triangles (3 per element)
  std::vector<double> xyz;      // nv x 3
  std::vector<int32_t> conn;    // ne x (dim + 1), 0-based
  std::vector<int32_t> labels;  // ne material labels (0 if absent)
};

struct SparseCoo {
  int64_t nrows = 0, ncols = 0;
  std::vector<int32_t> rows, cols;
  std::vector<double> vals;
};

// return 0 on success, non-zero (with err set) otherwise
int read_tetgen(const std::string& base, MeshData& m, std::string& err);
int read_trimesh(const std::string& fname, MeshData& m, std::string& err);
int read_matlab_sparse(const std::string& fname, SparseCoo& A, std::string& err);  // 0 / 1 / -1
int read_matlab_array(const std::string& fname, std::vector<double>& out, std::string& err);  // 0 / -1
int write_matlab_array(const std::string& fname, const double* x, int64_t n, std::string& err);
int write_vtk(const std::string& fname, int dim, int64_t nv, const double* xyz, int64_t ne, const int32_t* conn,
              const double* values, int64_t nvalues, std::string& err);

}  // namespace fem
