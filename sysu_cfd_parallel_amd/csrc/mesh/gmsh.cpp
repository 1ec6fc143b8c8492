// Mesh files in the two formats the reference reads:
//  * gmsh 2 ASCII (RootProcess::read, RootStructure.cpp:428-487): skip to
//    $Nodes, n vertices "id x y z", then $EndNodes $Elements, m elements
//    "id type ntag tag... v1 v2 ..."; vertex references are file id - 1;
//  * the binary form written by postCycas.py and read by RootProcess::readBin
//    (RootStructure.cpp:61-116): size_t n, n x {int id, double x, y, z},
//    size_t m, m x {int id, int type, int ntag, int tag[ntag], int v[nv]}.
// The first tag of an element is its physical tag (region id).  Vertices no
// element references are dropped and the rest renumbered in increasing
// order, as the reference's vertex pool does (RootStructure.cpp:930-968).
#include <cstdio>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "gmsh.hpp"

namespace s355 {

namespace {
struct RawMesh {
  std::vector<double> xyz;
  std::vector<int> type, tag, vptr{0}, vidx;  // vidx: file id - 1
};

ElementStream compact(RawMesh& raw) {
  const long nv = static_cast<long>(raw.xyz.size() / 3);
  std::vector<int> used(nv, 0);
  for (int v : raw.vidx) {
    if (v < 0 || v >= nv) throw std::runtime_error("mesh file: element vertex " + std::to_string(v + 1) + " out of range");
    used[v] = 1;
  }
  std::vector<int> local(nv, -1);
  ElementStream es;
  int next = 0;
  for (long v = 0; v < nv; ++v) {
    if (!used[v]) continue;
    local[v] = next++;
    es.xyz.insert(es.xyz.end(), raw.xyz.begin() + 3 * v, raw.xyz.begin() + 3 * v + 3);
  }
  es.type = std::move(raw.type);
  es.tag = std::move(raw.tag);
  es.vptr = std::move(raw.vptr);
  es.vidx.resize(raw.vidx.size());
  for (size_t k = 0; k < raw.vidx.size(); ++k) es.vidx[k] = local[raw.vidx[k]];
  return es;
}

void check_type(int type) {
  if (gmsh_nodes(type) < 0) throw std::runtime_error("mesh file: unsupported element type " + std::to_string(type));
}
}  // namespace

ElementStream read_gmsh_ascii(const std::string& path) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("cannot open mesh file " + path);
  std::string tok;
  do {
    if (!(in >> tok)) throw std::runtime_error("mesh file: no $Nodes section");
  } while (tok != "$Nodes");
  long nv = 0;
  if (!(in >> nv) || nv <= 0) throw std::runtime_error("mesh file: vertex count");
  RawMesh raw;
  raw.xyz.resize(3 * static_cast<size_t>(nv));
  for (long i = 0; i < nv; ++i) {
    long id;
    if (!(in >> id >> raw.xyz[3 * i] >> raw.xyz[3 * i + 1] >> raw.xyz[3 * i + 2]))
      throw std::runtime_error("mesh file: truncated $Nodes");
  }
  in >> tok >> tok;  // $EndNodes $Elements
  long ne = 0;
  if (!(in >> ne) || ne <= 0) throw std::runtime_error("mesh file: element count");
  for (long e = 0; e < ne; ++e) {
    long id;
    int type, ntag;
    if (!(in >> id >> type >> ntag) || ntag < 1) throw std::runtime_error("mesh file: truncated $Elements");
    check_type(type);
    int tag0 = 0;
    for (int t = 0; t < ntag; ++t) {
      int tg;
      in >> tg;
      if (t == 0) tag0 = tg;
    }
    for (int k = 0; k < gmsh_nodes(type); ++k) {
      long v;
      if (!(in >> v)) throw std::runtime_error("mesh file: truncated element");
      raw.vidx.push_back(static_cast<int>(v - 1));
    }
    raw.type.push_back(type);
    raw.tag.push_back(tag0);
    raw.vptr.push_back(static_cast<int>(raw.vidx.size()));
  }
  return compact(raw);
}

ElementStream read_gmsh_binary(const std::string& path) {
  std::ifstream in(path, std::ios::binary);
  if (!in) throw std::runtime_error("cannot open mesh file " + path);
  auto rd = [&](void* p, size_t n) {
    if (!in.read(static_cast<char*>(p), static_cast<std::streamsize>(n)))
      throw std::runtime_error("binary mesh file: truncated");
  };
  size_t nv = 0;
  rd(&nv, sizeof nv);
  if (nv == 0 || nv > (size_t(1) << 31)) throw std::runtime_error("binary mesh file: vertex count");
  RawMesh raw;
  raw.xyz.resize(3 * nv);
  for (size_t i = 0; i < nv; ++i) {
    int id;
    rd(&id, sizeof id);
    rd(&raw.xyz[3 * i], 3 * sizeof(double));
  }
  size_t ne = 0;
  rd(&ne, sizeof ne);
  if (ne == 0 || ne > (size_t(1) << 31)) throw std::runtime_error("binary mesh file: element count");
  for (size_t e = 0; e < ne; ++e) {
    int head[3];  // id, type, ntag
    rd(head, sizeof head);
    check_type(head[1]);
    if (head[2] < 1 || head[2] > 64) throw std::runtime_error("binary mesh file: tag count");
    std::vector<int> tags(head[2]);
    rd(tags.data(), sizeof(int) * tags.size());
    const int nn = gmsh_nodes(head[1]);
    int v[8];
    rd(v, sizeof(int) * nn);
    for (int k = 0; k < nn; ++k) raw.vidx.push_back(v[k] - 1);
    raw.type.push_back(head[1]);
    raw.tag.push_back(tags[0]);
    raw.vptr.push_back(static_cast<int>(raw.vidx.size()));
  }
  return compact(raw);
}

void write_gmsh_ascii(const ElementStream& es, const std::string& path) {
  FILE* f = std::fopen(path.c_str(), "w");
  if (!f) throw std::runtime_error("cannot write mesh file " + path);
  std::fprintf(f, "$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n%d\n", es.n_vert());
  for (int i = 0; i < es.n_vert(); ++i)
    std::fprintf(f, "%d %.17g %.17g %.17g\n", i + 1, es.xyz[3 * i], es.xyz[3 * i + 1], es.xyz[3 * i + 2]);
  std::fprintf(f, "$EndNodes\n$Elements\n%d\n", es.n_elem());
  for (int e = 0; e < es.n_elem(); ++e) {
    std::fprintf(f, "%d %d 2 %d %d", e + 1, es.type[e], es.tag[e], es.tag[e]);
    for (int k = es.vptr[e]; k < es.vptr[e + 1]; ++k) std::fprintf(f, " %d", es.vidx[k] + 1);
    std::fputc('\n', f);
  }
  std::fprintf(f, "$EndElements\n");
  if (std::fclose(f) != 0) throw std::runtime_error("error writing " + path);
}

void write_gmsh_binary(const ElementStream& es, const std::string& path) {
  std::ofstream out(path, std::ios::binary);
  if (!out) throw std::runtime_error("cannot write mesh file " + path);
  auto wr = [&](const void* p, size_t n) { out.write(static_cast<const char*>(p), static_cast<std::streamsize>(n)); };
  const size_t nv = es.n_vert(), ne = es.n_elem();
  wr(&nv, sizeof nv);
  for (size_t i = 0; i < nv; ++i) {
    const int id = static_cast<int>(i) + 1;
    wr(&id, sizeof id);
    wr(&es.xyz[3 * i], 3 * sizeof(double));
  }
  wr(&ne, sizeof ne);
  for (size_t e = 0; e < ne; ++e) {
    const int head[5] = {static_cast<int>(e) + 1, es.type[e], 2, es.tag[e], es.tag[e]};
    wr(head, sizeof head);
    for (int k = es.vptr[e]; k < es.vptr[e + 1]; ++k) {
      const int v = es.vidx[k] + 1;
      wr(&v, sizeof v);
    }
  }
  if (!out) throw std::runtime_error("error writing " + path);
}

}  // namespace s355
