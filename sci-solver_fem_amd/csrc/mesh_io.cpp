// Host-side file formats: TetGen .node/.ele, ASCII PLY / OFF triangle meshes,
// MATLAB v5 (uncompressed) sparse and dense arrays, legacy VTK output.
//
// Each reader follows the byte/line layout the reference accepts:
//   TetMesh::read                 src/core/cuda/tetmesh.cu:251-376
//   TriMesh::read (PLY, OFF)      src/core/aggmis/cuda/TriMesh_io.cu:146-257,259-502,692-708
//   readMatlabSparseMatrix        src/FEMSolver.cu:177-356
//   readMatlabArray               src/FEMSolver.cu:358-449
//   writeMatlabArray              src/FEMSolver.cu:451-517
//   writeVTK                      src/FEMSolver.cu:519-582
#include "mesh_io.hpp"

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace fem {

namespace {

bool skip_line(const std::string& line) { return line.empty() || line[0] == '#'; }

// Initial capacity for a list whose length comes from a file header: the
// vectors grow as lines are read, so a corrupt or absurd count fails with a
// "expected N" message instead of a huge allocation up front.
size_t header_reserve(int64_t n) { return (size_t)std::min<int64_t>(std::max<int64_t>(n, 0), (int64_t)1 << 20); }

}  // namespace

int read_tetgen(const std::string& base, MeshData& m, std::string& err) {
  std::ifstream nf(base + ".node"), ef(base + ".ele");
  if (!nf.is_open() || !ef.is_open()) {
    err = "node or ele file open failed: " + base;
    return 2;
  }
  std::string line;
  int nv = -1, t0, t1, t2;
  while (std::getline(nf, line)) {
    if (skip_line(line)) continue;
    if (std::sscanf(line.c_str(), "%d %d %d %d", &nv, &t0, &t1, &t2) != 4 || nv < 0) {
      err = "Bad Node file";
      return 2;
    }
    break;
  }
  if (nv < 0) {
    err = "Bad Node file";
    return 2;
  }
  m.dim = 3;
  m.xyz.clear();
  m.xyz.reserve(3 * header_reserve(nv));
  int64_t i = 0;
  while (i < nv && std::getline(nf, line)) {
    if (skip_line(line)) continue;
    int id;
    float x, y, z;  // the reference parses coordinates as float
    if (std::sscanf(line.c_str(), "%d %f %f %f", &id, &x, &y, &z) != 4) {
      err = "Bad Node file, line # " + std::to_string(i);
      return 2;
    }
    m.xyz.push_back(x);
    m.xyz.push_back(y);
    m.xyz.push_back(z);
    ++i;
  }
  if (i != nv) {
    err = "Bad Node file: expected " + std::to_string(nv) + " vertices";
    return 2;
  }
  int ne = -1, haslabel = 0;
  while (std::getline(ef, line)) {
    if (skip_line(line)) continue;
    if (std::sscanf(line.c_str(), "%d %d %d", &ne, &t0, &haslabel) != 3 || ne < 0) {
      err = "Bad Ele file";
      return 2;
    }
    break;
  }
  if (ne < 0) {
    err = "Bad Ele file";
    return 2;
  }
  m.conn.clear();
  m.labels.clear();
  m.conn.reserve(4 * header_reserve(ne));
  m.labels.reserve(header_reserve(ne));
  bool zero_based = false;
  int64_t e = 0;
  while (e < ne && std::getline(ef, line)) {
    if (skip_line(line)) continue;
    int id, v[4], mat = 0;
    int got = haslabel == 0 ? std::sscanf(line.c_str(), "%d %d %d %d %d", &id, &v[0], &v[1], &v[2], &v[3])
                            : std::sscanf(line.c_str(), "%d %d %d %d %d %d", &id, &v[0], &v[1], &v[2], &v[3], &mat);
    if (got != (haslabel == 0 ? 5 : 6)) {
      err = "Bad Ele file, line # " + std::to_string(e);
      return 2;
    }
    for (int a = 0; a < 4; ++a) {
      m.conn.push_back(v[a]);
      if (v[a] == 0) zero_based = true;
    }
    m.labels.push_back(mat);
    ++e;
  }
  if (e != ne) {
    err = "Bad Ele file: expected " + std::to_string(ne) + " elements";
    return 2;
  }
  if (!zero_based)
    for (auto& c : m.conn) --c;
  // second pass of the reference: a minimum index of 1 is shifted again
  int mn = 0x7fffffff;
  for (int c : m.conn) mn = std::min(mn, c);
  if (mn == 1)
    for (auto& c : m.conn) --c;
  for (int c : m.conn)
    if (c < 0 || c >= nv) {
      err = "element references a vertex outside the node file";
      return 2;
    }
  return 0;
}

namespace {

bool next_token_line(FILE* f, std::string& out) {
  char buf[4096];
  while (std::fgets(buf, sizeof(buf), f)) {
    std::string s(buf);
    size_t a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) continue;
    if (s[a] == '#') continue;
    out = s;
    return true;
  }
  return false;
}

int read_faces_ascii(FILE* f, int64_t nfaces, int skip_before, MeshData& m, std::string& err) {
  std::string line;
  for (int64_t k = 0; k < nfaces; ++k) {
    if (!next_token_line(f, line)) {
      err = "unexpected end of face list";
      return 2;
    }
    std::istringstream is(line);
    double dummy;
    for (int q = 0; q < skip_before; ++q) is >> dummy;
    int cnt;
    if (!(is >> cnt) || cnt < 3 || cnt > 65536) {
      err = "bad face line";
      return 2;
    }
    std::vector<int> v(cnt);
    for (int q = 0; q < cnt; ++q)
      if (!(is >> v[q])) {
        err = "bad face line";
        return 2;
      }
    for (int q = 1; q + 1 < cnt; ++q) {  // fan triangulation of polygons
      m.conn.push_back(v[0]);
      m.conn.push_back(v[q]);
      m.conn.push_back(v[q + 1]);
    }
  }
  return 0;
}

}  // namespace

int read_trimesh(const std::string& fname, MeshData& m, std::string& err) {
  FILE* f = std::fopen(fname.c_str(), "rb");
  if (!f) {
    err = "could not open mesh file: " + fname;
    return 2;
  }
  m.dim = 2;
  m.xyz.clear();
  m.conn.clear();
  std::string line;
  if (!next_token_line(f, line)) {
    std::fclose(f);
    err = "Can't read header";
    return 2;
  }
  int rc = 0;
  if (line.compare(0, 3, "ply") == 0) {
    int64_t nverts = 0, nfaces = 0;
    int nprops = 0, ix = -1, iy = -1, iz = -1;
    bool in_vertex = false, ascii = false;
    while (next_token_line(f, line)) {
      std::istringstream is(line);
      std::string w;
      is >> w;
      if (w == "format") {
        std::string fmt;
        is >> fmt;
        ascii = (fmt == "ascii");
      } else if (w == "element") {
        std::string what;
        int64_t cnt;
        is >> what >> cnt;
        in_vertex = (what == "vertex");
        if (in_vertex) nverts = cnt;
        if (what == "face") nfaces = cnt;
      } else if (w == "property" && in_vertex) {
        std::string type, name;
        is >> type >> name;
        if (name == "x") ix = nprops;
        if (name == "y") iy = nprops;
        if (name == "z") iz = nprops;
        ++nprops;
      } else if (w == "end_header") {
        break;
      }
    }
    if (!ascii || ix < 0 || iy < 0) {
      std::fclose(f);
      err = "only ASCII PLY files with x/y vertex properties are supported";
      return 2;
    }
    if (nverts < 0 || nfaces < 0) {
      std::fclose(f);
      err = "negative element count in PLY header";
      return 2;
    }
    m.xyz.reserve(3 * header_reserve(nverts));
    for (int64_t v = 0; v < nverts; ++v) {
      if (!next_token_line(f, line)) {
        std::fclose(f);
        err = "unexpected end of vertex list";
        return 2;
      }
      std::istringstream is(line);
      std::vector<float> p(nprops, 0.f);  // TriMesh stores float points
      for (int q = 0; q < nprops; ++q) is >> p[q];
      m.xyz.push_back(p[ix]);
      m.xyz.push_back(p[iy]);
      m.xyz.push_back(iz >= 0 ? p[iz] : 0.f);
    }
    rc = read_faces_ascii(f, nfaces, 0, m, err);
  } else if (line.compare(0, 3, "OFF") == 0) {
    std::string counts;
    size_t rest = line.find_first_not_of(" \t\r\n", 3);
    if (rest != std::string::npos) counts = line.substr(rest);
    else if (!next_token_line(f, counts)) counts.clear();
    int64_t nverts = 0, nfaces = 0;
    if (std::sscanf(counts.c_str(), "%ld %ld", &nverts, &nfaces) < 2) {
      std::fclose(f);
      err = "bad OFF header";
      return 2;
    }
    if (nverts < 0 || nfaces < 0) {
      std::fclose(f);
      err = "negative count in OFF header";
      return 2;
    }
    m.xyz.reserve(3 * header_reserve(nverts));
    for (int64_t v = 0; v < nverts; ++v) {
      if (!next_token_line(f, line)) {
        std::fclose(f);
        err = "unexpected end of vertex list";
        return 2;
      }
      float x = 0, y = 0, z = 0;
      std::sscanf(line.c_str(), "%f %f %f", &x, &y, &z);
      m.xyz.push_back(x);
      m.xyz.push_back(y);
      m.xyz.push_back(z);
    }
    rc = read_faces_ascii(f, nfaces, 0, m, err);
  } else {
    err = "Unknown file type (ASCII PLY and OFF are supported)";
    rc = 2;
  }
  std::fclose(f);
  if (rc) return rc;
  int64_t nv = (int64_t)m.xyz.size() / 3;
  for (int c : m.conn)
    if (c < 0 || c >= nv) {
      err = "face references a vertex outside the vertex list";
      return 2;
    }
  m.labels.assign(m.conn.size() / 3, 0);
  return 0;
}

// ---- MAT v5 -----------------------------------------------------------------

namespace {

struct Reader {
  std::ifstream in;
  bool ok = true;
  template <typename T>
  T get() {
    T v{};
    in.read(reinterpret_cast<char*>(&v), sizeof(T));
    if (!in) ok = false;
    return v;
  }
  void skip(int64_t n) {
    if (n <= 0) return;
    in.seekg(n, std::ios::cur);
    if (!in) ok = false;
  }
  void bytes(void* p, int64_t n) {
    in.read(reinterpret_cast<char*>(p), n);
    if (!in) ok = false;
  }
  // bytes left after the read position (0 once a read has failed)
  int64_t remaining() {
    if (!ok) return 0;
    const std::streampos here = in.tellg();
    in.seekg(0, std::ios::end);
    const std::streampos end = in.tellg();
    in.seekg(here);
    if (!in || here < 0 || end < here) {
      ok = false;
      return 0;
    }
    return (int64_t)(end - here);
  }
  // a data element of n bytes must be non-negative and fit in the file
  bool fits(int64_t n) { return n >= 0 && n <= remaining(); }
};

}  // namespace

int read_matlab_sparse(const std::string& fname, SparseCoo& A, std::string& err) {
  Reader r;
  r.in.open(fname.c_str(), std::ios::binary);
  if (!r.in.is_open()) {
    err = "could not open file: " + fname;
    return 1;
  }
  r.skip(128);
  int32_t type = r.get<int32_t>();
  if (type == 15) {
    err = "Compression not supported. Save matlab data with '-v6' option.";
    return 1;
  }
  if (type != 14) {
    err = fname + " is not a matlab matrix.";
    return 1;
  }
  (void)r.get<uint32_t>();  // data size
  type = r.get<int32_t>();
  if (type != 6 && type != 5) {
    err = "Invalid type for sparse matrix. Must be 32bit.";
    return 1;
  }
  (void)r.get<int32_t>();  // flags size
  uint32_t mclass = r.get<uint32_t>() & 0xFFu;
  if (mclass != 5) {
    err = "This is not a sparse matrix file.";
    return 1;
  }
  (void)r.get<uint32_t>();  // nzmax
  type = r.get<int32_t>();
  int32_t bpe = r.get<int32_t>();
  if ((type != 6 && type != 5) || bpe != 8) {
    err = "Matrix of wrong dimension type or # of dimensions.";
    return 1;
  }
  int32_t xdim = r.get<int32_t>(), ydim = r.get<int32_t>();
  uint32_t name_type = r.get<uint16_t>();
  uint32_t name_len = r.get<uint16_t>();
  uint32_t align = 4;
  if (name_len == 0) {
    name_len = r.get<uint32_t>();
    align = 8;
  }
  if (name_type != 1 && name_type != 2) {
    err = "Invalid variable type (" + std::to_string(name_type) + ") for array name characters (Must be 8-bit).";
    return -1;
  }
  if (name_len % align) name_len += align - name_len % align;
  r.skip(name_len);
  // row indices
  type = r.get<int32_t>();
  if (type != 6 && type != 5) {
    err = "Invalid type row index for sparse matrix. Must be 32bit.";
    return 1;
  }
  bpe = r.get<int32_t>();
  if (!r.fits(bpe)) {
    err = "truncated sparse matrix file";
    return 1;
  }
  std::vector<int32_t> ir(bpe / 4);
  r.bytes(ir.data(), (int64_t)ir.size() * 4);
  r.skip(bpe % 8);
  // column pointers
  type = r.get<int32_t>();
  if (type != 6 && type != 5) {
    err = "Invalid column index type for sparse matrix. Must be 32bit.";
    return 1;
  }
  bpe = r.get<int32_t>();
  if (!r.fits(bpe)) {
    err = "truncated sparse matrix file";
    return 1;
  }
  std::vector<int32_t> jc(bpe / 4);
  r.bytes(jc.data(), (int64_t)jc.size() * 4);
  r.skip(bpe % 8);
  // values
  type = r.get<int32_t>();
  if (type != 9) {
    err = "Invalid value for sparse matrix. Must be double float.";
    return 1;
  }
  bpe = r.get<int32_t>();
  if (!r.fits(bpe)) {
    err = "truncated sparse matrix file";
    return 1;
  }
  std::vector<double> pr(bpe / 8);
  r.bytes(pr.data(), (int64_t)pr.size() * 8);
  if (!r.ok || xdim < 0 || ydim < 0 || (int64_t)jc.size() < (int64_t)ydim + 1) {
    err = "truncated sparse matrix file";
    return 1;
  }
  A.nrows = xdim;
  A.ncols = ydim;
  A.rows.clear();
  A.cols.clear();
  A.vals.clear();
  for (int32_t c = 0; c < ydim; ++c)
    for (int32_t p = jc[c]; p < jc[c + 1]; ++p) {
      if (p < 0 || p >= (int32_t)ir.size() || p >= (int32_t)pr.size() || ir[p] < 0 || ir[p] >= xdim) {
        err = "sparse matrix index out of range";
        return 1;
      }
      A.rows.push_back(ir[p]);
      A.cols.push_back(c);
      A.vals.push_back(pr[p]);
    }
  return 0;
}

int read_matlab_array(const std::string& fname, std::vector<double>& out, std::string& err) {
  Reader r;
  r.in.open(fname.c_str(), std::ios::binary);
  if (!r.in.is_open()) {
    err = "could not open file: " + fname;
    return -1;
  }
  r.skip(128);
  int32_t type = r.get<int32_t>();
  if (type == 15) {
    err = "Compression not supported. Save matlab data with '-v6' option.";
    return -1;
  }
  if (type != 14) {
    err = fname + " is not a matlab matrix.";
    return -1;
  }
  (void)r.get<uint32_t>();
  type = r.get<int32_t>();
  if (type != 6) {
    err = "Invalid type for normal matrix. Must be double precision.";
    return -1;
  }
  (void)r.get<int32_t>();
  uint32_t mclass = r.get<uint32_t>() & 0xFFu;
  if (mclass == 5) {
    err = "This import routine is not for a sparse matrix file.";
    return -1;
  }
  (void)r.get<uint32_t>();
  type = r.get<int32_t>();
  int32_t bpe = r.get<int32_t>();
  if ((type != 6 && type != 5) || bpe != 8) {
    err = "Matrix of wrong dimension type or # of dimensions.";
    return -1;
  }
  (void)r.get<int32_t>();
  (void)r.get<int32_t>();
  uint32_t name_type = r.get<uint16_t>();
  if (name_type != 1 && name_type != 2) {
    err = "Invalid variable type (" + std::to_string(name_type) + ") for array name characters (Must be 8-bit).";
    return -1;
  }
  uint32_t name_len = r.get<uint16_t>();
  if (name_len % 4) name_len += 4 - name_len % 4;
  r.skip(name_len);
  type = r.get<int32_t>();
  if (type != 9) {
    err = "Matrix data type must be miDOUBLE (type is " + std::to_string(type) + ").";
    return -1;
  }
  uint32_t len = r.get<uint32_t>();
  if (!r.fits((int64_t)len)) {
    err = "truncated array file";
    return -1;
  }
  out.assign(len / 8, 0.0);
  r.bytes(out.data(), (int64_t)out.size() * 8);
  if (!r.ok) {
    err = "truncated array file";
    return -1;
  }
  return 0;
}

int write_matlab_array(const std::string& fname, const double* x, int64_t n, std::string& err) {
  std::ofstream f(fname.c_str(), std::ios::out | std::ios::binary);
  if (!f.is_open()) {
    err = "could not open file for writing: " + fname;
    return 1;
  }
  std::string desc = "MATLAB 5.0 MAT-file, Platform: GLNXA64, Created by SCI-Solver_FEM.";
  desc.resize(116, ' ');
  f.write(desc.data(), 116);
  char zeros[8] = {0};
  f.write(zeros, 8);
  int16_t version = 0x0100;
  f.write(reinterpret_cast<const char*>(&version), 2);
  f.write("IM", 2);
  auto put32 = [&](int32_t v) { f.write(reinterpret_cast<const char*>(&v), 4); };
  auto put16 = [&](int16_t v) { f.write(reinterpret_cast<const char*>(&v), 2); };
  put32(14);
  const int32_t total = 8 + 8 + 8 + 8 + 8 + 8 + (int32_t)(8 * n);  // flags, dims, name, data header + data
  put32(total);
  put32(6);
  put32(8);
  put32(6);  // class double
  f.write(zeros, 4);
  put32(5);
  put32(8);
  put32((int32_t)n);
  put32(1);
  put16(1);
  put16(3);
  const char name[4] = {'x', '_', 'h', '\0'};
  f.write(name, 4);
  put32(9);
  put32((int32_t)(n * 8));
  f.write(reinterpret_cast<const char*>(x), n * 8);
  if (!f) {
    err = "write failed: " + fname;
    return 1;
  }
  return 0;
}

int write_vtk(const std::string& fname, int dim, int64_t nv, const double* xyz, int64_t ne, const int32_t* conn,
              const double* values, int64_t nvalues, std::string& err) {
  FILE* f = std::fopen((fname + ".vtk").c_str(), "w+");
  if (!f) {
    err = "could not open " + fname + ".vtk";
    return 1;
  }
  std::fprintf(f, "# vtk DataFile Version 3.0\nvtk output\nASCII\nDATASET UNSTRUCTURED_GRID\n");
  std::fprintf(f, "POINTS %d float\n", (int)nv);
  for (int64_t i = 0; i < nv; ++i) std::fprintf(f, "%.12f %.12f %.12f\n", xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
  const int npe = dim + 1;
  std::fprintf(f, "CELLS %d %d\n", (int)ne, (int)(ne * (npe + 1)));
  for (int64_t e = 0; e < ne; ++e) {
    std::fprintf(f, "%d", npe);
    for (int a = 0; a < npe; ++a) std::fprintf(f, " %d", conn[npe * e + a]);
    std::fprintf(f, "\n");
  }
  std::fprintf(f, "CELL_TYPES %d\n", (int)ne);
  for (int64_t e = 0; e < ne; ++e) std::fprintf(f, dim == 3 ? "10\n" : "5\n");
  std::fprintf(f, "POINT_DATA %d\nSCALARS traveltime float 1\nLOOKUP_TABLE default\n", (int)nv);
  if (dim == 3) {
    for (int64_t i = 0; i < nvalues; ++i) std::fprintf(f, "%.12f\n ", values[i]);
  } else {
    for (int64_t i = 0; i < nv && i < nvalues; ++i) std::fprintf(f, "%.12f\n", (double)(float)values[i]);
  }
  std::fclose(f);
  return 0;
}

}  // namespace fem
