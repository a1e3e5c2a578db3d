// Usage contract of the reference's src/examples/example2.cu (triangle meshes):
//   example2 [-v] [-i mesh.ply|mesh.off] [-A matrix.mat] [-b rhs.mat]
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "FEMSolver.h"

int main(int argc, char** argv) {
  std::string Aname, bName, fname = "../src/test/test_data/sphere_290verts.ply";
  bool verbose = false;
  for (int i = 0; i < argc; i++) {
    if (std::strcmp(argv[i], "-v") == 0) {
      verbose = true;
    } else if (std::strcmp(argv[i], "-i") == 0 && i + 1 < argc) {
      fname = argv[++i];
    } else if (std::strcmp(argv[i], "-b") == 0 && i + 1 < argc) {
      bName = argv[++i];
    } else if (std::strcmp(argv[i], "-A") == 0 && i + 1 < argc) {
      Aname = argv[++i];
    }
  }
  try {
    FEMSolver cfg(fname, false, verbose);
    if (!Aname.empty() && cfg.readMatlabSparseMatrix(Aname) != 0)
      std::cerr << "Failed to read in A matrix: " << Aname << std::endl;
    Vector_h_CG b_h(cfg.getMatrixRows(), 1.0);
    if (!bName.empty() && cfg.readMatlabArray(bName, &b_h) != 0)
      std::cerr << "Failed to read in b array: " << bName << std::endl;
    Vector_h_CG x_h(cfg.getMatrixRows(), 0.0);
    cfg.solveFEM(&x_h, &b_h);
    if (cfg.writeMatlabArray("output.mat", x_h)) std::cerr << "failed to write matlab file." << std::endl;
    std::vector<double> vals(x_h.begin(), x_h.end());
    // output base name: the mesh file name without its folders and extension
    // (example2 writes <name>.vtk next to the working directory)
    std::string outname = cfg.filename_;
    const size_t slash = outname.find_last_of("/\\");
    if (slash != std::string::npos) outname = outname.substr(slash + 1);
    const size_t dot = outname.find_last_of('.');
    if (dot != std::string::npos) outname = outname.substr(0, dot);
    cfg.writeVTK(vals, outname);
  } catch (const std::exception& e) {
    std::cerr << e.what() << std::endl;
    return 1;
  }
  return 0;
}
