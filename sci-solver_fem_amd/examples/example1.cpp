// Usage contract of the reference's src/examples/example1.cu (tet meshes):
//   example1 [-v] [-i mesh_base] [-A matrix.mat] [-b rhs.mat]
// builds FEMSolver from <mesh_base>.node/.ele, optionally replaces A and b
// from MATLAB files, solves, writes output.mat and <mesh name>.vtk.
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "FEMSolver.h"

int main(int argc, char** argv) {
  std::string Aname, bName, fname = "../src/test/test_data/CubeMesh_size256step16";
  bool verbose = false;
  bool pcg = false;
  for (int i = 0; i < argc; i++) {
    if (std::strcmp(argv[i], "-v") == 0) {
      verbose = true;
    } else if (std::strcmp(argv[i], "-i") == 0 && i + 1 < argc) {
      fname = argv[++i];
    } else if (std::strcmp(argv[i], "-b") == 0 && i + 1 < argc) {
      bName = argv[++i];
    } else if (std::strcmp(argv[i], "-A") == 0 && i + 1 < argc) {
      Aname = argv[++i];
    } else if (std::strcmp(argv[i], "-pcg") == 0) {
      pcg = true;
    }
  }
  try {
    FEMSolver cfg(fname, true, verbose);
    if (pcg) {
      cfg.solverType_ = 1;
      cfg.tolerance_ = 1e-10;
    }
    if (!Aname.empty() && cfg.readMatlabSparseMatrix(Aname) != 0)
      std::cerr << "Failed to read in A matrix: " << Aname << std::endl;
    Vector_h_CG b_h(cfg.getMatrixRows(), 1.0);
    if (!bName.empty() && cfg.readMatlabArray(bName, &b_h) != 0)
      std::cerr << "Failed to read in b array: " << bName << std::endl;
    Vector_h_CG x_h(cfg.getMatrixRows(), 0.0);
    cfg.solveFEM(&x_h, &b_h);
    if (cfg.writeMatlabArray("output.mat", x_h)) std::cerr << "failed to write matlab file." << std::endl;
    std::vector<double> vals(x_h.begin(), x_h.end());
    size_t pos = cfg.filename_.find_last_of("/");
    std::string outname = (pos == std::string::npos) ? cfg.filename_ : cfg.filename_.substr(pos + 1);
    cfg.writeVTK(vals, outname);
    std::cout << "iterations " << cfg.lastSolveInfo().iterations << " residual " << cfg.lastSolveInfo().rel_residual
              << std::endl;
  } catch (const std::exception& e) {
    std::cerr << e.what() << std::endl;
    return 1;
  }
  return 0;
}
