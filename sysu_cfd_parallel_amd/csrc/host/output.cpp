// Result and restart files, byte for byte in the reference's layouts:
//  * Tecplot ASCII (Output2Tecplot, navierIO.cpp:161-246): a variables line,
//    one FEBRICK zone in BLOCK packing with the vertex coordinates NODAL and
//    p u v w ro T Ma mu (te ed) CELLCENTERED, each array written by
//    OutArray2File (navierIO.cpp:2-16) as "%19.5f " per value with a newline
//    after every fifth and one more at the end, then one line of the cell's
//    8 vertex ids (1-based) per cell, each line followed by a space;
//  * tec/data (Output2TecplotBinary, navierIO.cpp:132-159): int rank, int
//    nVar, then P U V W R T (TE ED) as Ncel doubles each;
//  * tec/geo (OutputGridBinary, navierIO.cpp:29-56): int rank, nProcess,
//    Ncel, Nvrt; Nvrt x 3 doubles; Ncel x 8 vertex ids (1-based);
//  * tec/grid.dat (OutputGrid, navierIO.cpp:58-122): fepoint zone with a
//    fourth per-vertex value naming the neighbour partition (0 on one rank);
//  * res.sav (WriteBackupFile / ReadBackupFile, navierIO.cpp:370-403): per
//    cell "R U V W P T R VisLam" and, with k-epsilon, "VisTur TE ED" lines in
//    the stream's default format (6 significant digits, as %g);
//  * res.vtk (Output2VTK, navierIO.cpp:287-365): legacy ASCII unstructured
//    grid of 8-node cells (type 12, 0-based ids), the velocity vector and the
//    scalars P ro T mach mu (te ed) Elemtype as floats, stream format.
// One process writes the whole file (rank 0 of 1 partition).  Values of
// 1e13 and above overflow the 19-character field: the reference's buffer
// arithmetic then loses characters, here they are written in full.
#include "output.hpp"

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <vector>

#include "../mesh/parallel.hpp"

namespace s355host {

namespace {

// long enough for four doubles in %f (at most 309 integer digits each)
constexpr int kBuf = 2048;

class File {
 public:
  File(const std::string& path, const char* mode) : path_(path), f_(std::fopen(path.c_str(), mode)) {
    if (!f_) throw IoError("cannot open " + path + " for writing: " + std::strerror(errno));
  }
  ~File() {
    if (f_) std::fclose(f_);
  }
  void put(const void* p, size_t n) {
    if (n && std::fwrite(p, 1, n, f_) != n) throw IoError("error writing " + path_);
  }
  void put(const std::string& s) { put(s.data(), s.size()); }
  void close() {
    FILE* f = f_;
    f_ = nullptr;
    if (std::fclose(f) != 0) throw IoError("error writing " + path_);
  }

 private:
  std::string path_;
  FILE* f_;
};

// Formats items [0, n) in parallel chunks (chunk starts are multiples of
// `align` so per-item separators stay local) and writes them in order.
template <class Fmt>
void put_parallel(File& f, long n, long align, Fmt&& fmt) {
  const long chunk = align * ((1L << 16) / align + 1);
  const long nchunk = (n + chunk - 1) / chunk;
  const long batch = 64;  // chunks formatted together, bounding the memory held
  for (long c0 = 0; c0 < nchunk; c0 += batch) {
    const long c1 = std::min(nchunk, c0 + batch);
    std::vector<std::string> out(c1 - c0);
    s355::parallel_for(
        c1 - c0,
        [&](long c) {
          std::string& s = out[c];
          const long b = (c0 + c) * chunk, e = std::min(n, b + chunk);
          s.reserve(static_cast<size_t>(e - b) * 24);
          char buf[kBuf];
          for (long i = b; i < e; ++i) fmt(i, s, buf);
        },
        1);
    for (const auto& s : out) f.put(s);
  }
}

// OutArray2File: "%19.5f " per value, a newline after every fifth value,
// one more newline after the array.
void put_array(File& f, const double* a, long n, long stride) {
  put_parallel(f, n, 5, [&](long i, std::string& s, char* buf) {
    const int k = std::snprintf(buf, kBuf, "%19.5f ", a[i * stride]);
    s.append(buf, k);
    if (i % 5 == 4) s.push_back('\n');
  });
  f.put("\n", 1);
}

void check_view(const ResultView& r, bool fields) {
  if (r.n_cell < 1 || r.n_vert < 1 || !r.vert || !r.cell_vert) throw IoError("result: empty mesh view");
  if (fields && (!r.R || !r.U || !r.V || !r.W || !r.P || !r.T || !r.vislam))
    throw IoError("result: missing cell field");
  if (fields && r.turbulent && (!r.vistur || !r.te || !r.ed))
    throw IoError("result: k-epsilon output needs VisTur, TE and ED");
}

// "8 v0 ... v7 \n" per cell with the 0-based ids (Output2VTK's CELLS block)
void put_connectivity_raw(File& f, const ResultView& r) {
  put_parallel(f, r.n_cell, 1, [&](long i, std::string& s, char* buf) {
    const int* v = r.cell_vert + 8 * i;
    const int k = std::snprintf(buf, kBuf, "8 %d %d %d %d %d %d %d %d \n", v[0], v[1], v[2], v[3], v[4], v[5], v[6],
                                v[7]);
    s.append(buf, k);
  });
}

void put_connectivity(File& f, const ResultView& r, const char* fmt) {
  put_parallel(f, r.n_cell, 1, [&](long i, std::string& s, char* buf) {
    const int* v = r.cell_vert + 8 * i;
    const int k = std::snprintf(buf, kBuf, fmt, v[0] + 1, v[1] + 1, v[2] + 1, v[3] + 1, v[4] + 1, v[5] + 1, v[6] + 1,
                                v[7] + 1);
    s.append(buf, k);
  });
}

void tecplot_ascii(const ResultView& r, const std::string& path) {
  File f(path, "wb");
  std::string head = "variables= \"x\", \"y\", \"z\", \"p\", \"u\", \"v\", \"w\", \"ro\", \"T\",\"Ma\",\"mu\" ";
  if (r.turbulent) head += "\"te\", \"ed\"";
  head += "\n";
  f.put(head);
  const int nvar = r.turbulent ? 8 : 6;  // FlowField::nVar (BasicType.h:95-111)
  char zone[256];
  const int k = std::snprintf(zone, sizeof zone,
                              "ZONE N=%d, E=%d, VARLOCATION=([1-3]=NODAL,[4-%d]=CELLCENTERED) DATAPACKING=BLOCK, "
                              "ZONETYPE=FEBRICK\n",
                              r.n_vert, r.n_cell, 3 + nvar + 2);
  f.put(zone, k);
  for (int d = 0; d < 3; ++d) put_array(f, r.vert + d, r.n_vert, 3);
  const long n = r.n_cell;
  for (const double* a : {r.P, r.U, r.V, r.W, r.R, r.T}) put_array(f, a, n, 1);
  std::vector<double> tmp(n);
  // velocity magnitude in the Mach column for DensityModel 0
  s355::parallel_for(n, [&](long i) { tmp[i] = std::sqrt(r.U[i] * r.U[i] + r.V[i] * r.V[i] + r.W[i] * r.W[i]); });
  put_array(f, tmp.data(), n, 1);
  s355::parallel_for(n, [&](long i) { tmp[i] = r.vislam[i] + (r.vistur ? r.vistur[i] : 0.0); });
  put_array(f, tmp.data(), n, 1);
  if (r.turbulent) {
    put_array(f, r.te, n, 1);
    put_array(f, r.ed, n, 1);
  }
  put_connectivity(f, r, "%d %d %d %d %d %d %d %d\n ");
  f.close();
}

void tecplot_data(const ResultView& r, const std::string& path) {
  File f(path, "wb");
  const int head[2] = {0, r.turbulent ? 8 : 6};
  f.put(head, sizeof head);
  const size_t n = r.n_cell;
  for (const double* a : {r.P, r.U, r.V, r.W, r.R, r.T}) f.put(a, n * sizeof(double));
  if (r.turbulent) {
    f.put(r.te, n * sizeof(double));
    f.put(r.ed, n * sizeof(double));
  }
  f.close();
}

void grid_binary(const ResultView& r, const std::string& path) {
  File f(path, "wb");
  const int head[4] = {0, 1, r.n_cell, r.n_vert};
  f.put(head, sizeof head);
  f.put(r.vert, 3 * static_cast<size_t>(r.n_vert) * sizeof(double));
  std::vector<int> ids(8 * static_cast<size_t>(r.n_cell));
  s355::parallel_for(static_cast<long>(ids.size()), [&](long i) { ids[i] = r.cell_vert[i] + 1; });
  f.put(ids.data(), ids.size() * sizeof(int));
  f.close();
}

void grid_ascii(const ResultView& r, const std::string& path) {
  File f(path, "wb");
  f.put("title = \"cycas partition grid\"\nvariables=\"x\",\"y\",\"z\",\"p\"\n");
  char zone[128];
  const int k = std::snprintf(zone, sizeof zone, "zone n=%d, e=%d, f=fepoint, et=BRICK\n", r.n_vert, r.n_cell);
  f.put(zone, k);
  put_parallel(f, r.n_vert, 1, [&](long i, std::string& s, char* buf) {
    // fourth value: 1 + the neighbour partition sharing the vertex, 0 on one rank
    const int m = std::snprintf(buf, kBuf, "%f %f %f %f\n", r.vert[3 * i], r.vert[3 * i + 1], r.vert[3 * i + 2], 0.0);
    s.append(buf, m);
  });
  put_connectivity(f, r, "%8d %8d %8d %8d %8d %8d %8d %8d\n");
  f.close();
}

// Output2VTK; "%g" is the stream's default format, and the scalars are
// written as floats (outputVTKScalar, navierIO.cpp:18-25)
void vtk(const ResultView& r, const std::string& path) {
  File f(path, "wb");
  char head[256];
  int k = std::snprintf(head, sizeof head,
                        "# vtk DataFile Version 2.0\nNavier-Stokes solver result\nASCII\nDATASET UNSTRUCTURED_GRID\n"
                        "POINTS %d FLOAT \n",
                        r.n_vert);
  f.put(head, k);
  put_parallel(f, r.n_vert, 1, [&](long i, std::string& s, char* buf) {
    const int m = std::snprintf(buf, kBuf, "%g %g %g \n", r.vert[3 * i], r.vert[3 * i + 1], r.vert[3 * i + 2]);
    s.append(buf, m);
  });
  k = std::snprintf(head, sizeof head, "\nCELLS %d %ld\n", r.n_cell, 9L * r.n_cell);
  f.put(head, k);
  put_connectivity_raw(f, r);
  k = std::snprintf(head, sizeof head, "CELL_TYPES %d\n", r.n_cell);
  f.put(head, k);
  put_parallel(f, r.n_cell, 8, [&](long i, std::string& s, char*) {
    s.append("12 ");
    if ((i + 1) % 8 == 0) s.push_back('\n');
  });
  k = std::snprintf(head, sizeof head, "\nCELL_DATA %d\nVECTORS velocity FLOAT\n", r.n_cell);
  f.put(head, k);
  put_parallel(f, r.n_cell, 1, [&](long i, std::string& s, char* buf) {
    const int m = std::snprintf(buf, kBuf, "%g %g %g\n", r.U[i], r.V[i], r.W[i]);
    s.append(buf, m);
  });
  f.put("\n", 1);
  auto scalar = [&](const char* name, auto value) {
    char h[128];
    const int m = std::snprintf(h, sizeof h, "SCALARS %s FLOAT\nLOOKUP_TABLE default\n", name);
    f.put(h, m);
    put_parallel(f, r.n_cell, 1, [&](long i, std::string& s, char* buf) {
      const int q = std::snprintf(buf, kBuf, "%g\n", static_cast<double>(static_cast<float>(value(i))));
      s.append(buf, q);
    });
    f.put("\n", 1);
  };
  scalar("P", [&](long i) { return r.P[i]; });
  scalar("ro", [&](long i) { return r.R[i]; });
  scalar("T", [&](long i) { return r.T[i]; });
  scalar("mach", [&](long i) { return std::sqrt(r.U[i] * r.U[i] + r.V[i] * r.V[i] + r.W[i] * r.W[i]); });
  scalar("mu", [&](long i) { return r.vislam[i] + (r.vistur ? r.vistur[i] : 0.0); });
  if (r.turbulent) {
    scalar("te", [&](long i) { return r.te[i]; });
    scalar("ed", [&](long i) { return r.ed[i]; });
  }
  scalar("Elemtype", [](long) { return 2.0; });
  f.close();
}

void backup(const ResultView& r, const std::string& path) {
  File f(path, "wb");
  put_parallel(f, r.n_cell, 1, [&](long i, std::string& s, char* buf) {
    int m = std::snprintf(buf, kBuf, "%g %g %g %g %g %g %g %g\n", r.R[i], r.U[i], r.V[i], r.W[i], r.P[i], r.T[i],
                          r.R[i], r.vislam[i]);
    s.append(buf, m);
    if (r.turbulent) {
      m = std::snprintf(buf, kBuf, "%g %g %g\n", r.vistur[i], r.te[i], r.ed[i]);
      s.append(buf, m);
    }
  });
  f.close();
}

}  // namespace

void write_result(const ResultView& r, OutFormat fmt, const std::string& path) {
  const bool fields = fmt != OutFormat::GridBinary && fmt != OutFormat::GridAscii;
  check_view(r, fields);
  switch (fmt) {
    case OutFormat::TecplotAscii: tecplot_ascii(r, path); break;
    case OutFormat::TecplotData: tecplot_data(r, path); break;
    case OutFormat::GridBinary: grid_binary(r, path); break;
    case OutFormat::GridAscii: grid_ascii(r, path); break;
    case OutFormat::Backup: backup(r, path); break;
    case OutFormat::Vtk: vtk(r, path); break;
    default: throw IoError("unknown output format");
  }
}

void read_backup(const std::string& path, int n_cell, bool turbulent, const BackupArrays& a) {
  if (n_cell < 1) throw IoError("backup: no cells");
  if (!a.R || !a.U || !a.V || !a.W || !a.P || !a.T || !a.vislam || (turbulent && (!a.vistur || !a.te || !a.ed)))
    throw IoError("backup: missing output array");
  std::ifstream in(path, std::ios::binary);
  if (!in) throw IoError("cannot open backup file " + path);
  std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  const char* p = text.c_str();
  auto next = [&](double& v, long cell) {
    char* end = nullptr;
    v = std::strtod(p, &end);
    if (end == p) throw IoError("backup file " + path + ": truncated or malformed at cell " + std::to_string(cell));
    p = end;
  };
  for (long i = 0; i < n_cell; ++i) {
    // R U V W P T R VisLam: the second R is the one kept
    next(a.R[i], i);
    next(a.U[i], i);
    next(a.V[i], i);
    next(a.W[i], i);
    next(a.P[i], i);
    next(a.T[i], i);
    next(a.R[i], i);
    next(a.vislam[i], i);
    if (turbulent) {
      next(a.vistur[i], i);
      next(a.te[i], i);
      next(a.ed[i], i);
    }
  }
}

std::string monitor_header(const std::vector<MonitorRecord>& regions) {
  char buf[256];
  std::snprintf(buf, sizeof buf, "%20s || ", "time");
  std::string out = buf;
  for (const MonitorRecord& r : regions) {
    const bool wall = r.type1 == 1;
    std::snprintf(buf, sizeof buf, "%20s: %20d: %20s | %20s || ", wall ? "wall" : "fluid", r.rid,
                  wall ? "drag/lift force" : "Null", wall ? "Null" : "w energy");
    out += buf;
  }
  return out + "\n";
}

std::string monitor_line(double cur_time, const std::vector<MonitorRecord>& records) {
  char buf[256];
  std::snprintf(buf, sizeof buf, "%20e || ", cur_time);
  std::string out = buf;
  for (const MonitorRecord& r : records) {
    std::snprintf(buf, sizeof buf, "%20e, %20e, %20e | %20e || ", r.force[0], r.force[1], r.force[2], r.scalar);
    out += buf;
  }
  return out + "\n";
}

}  // namespace s355host
