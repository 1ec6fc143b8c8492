// Mesh ingest in the reference's numbering (drop-in order of cells, faces,
// boundary records and CSR rows), built for multi-million-cell meshes:
//   * element order: the single-partition std::sort of RootStructure.cpp:797-798
//     on equal keys, run on an index array so the permutation is libstdc++'s;
//   * vertices: ids compacted in ascending order (RootStructure.cpp:930-968);
//   * cells / boundary records: geometry.cpp:105-203 (8-node degenerate hexes);
//   * faces: geometry.cpp:222-358 with a hash keyed on the sorted distinct
//     vertex set instead of per-vertex face sets; boundary faces first,
//     interior faces numbered on first discovery in the pattern
//     (0123)(4567)(0154)(1265)(3267)(0374); faces with more than one repeated
//     vertex are skipped (geometry.cpp:303-306);
//   * metrics: geometry.cpp:391-752, same expressions in the same order;
//   * CSR: one row per cell, columns ascending, one diagonal slot.
// Extension beyond the reference (which aborts on such meshes): a hex quad
// covered by two triangles of split neighbours becomes those two triangles,
// and the cell's volume and centroid come from face triangles about the
// vertex mean.  The planarity check of geometry.cpp:604-610 counts warped
// quads instead of exiting.
#include "ingest.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>

namespace s355 {
namespace {

const int kPattern[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 1, 5, 4}, {1, 2, 6, 5}, {3, 2, 6, 7}, {0, 3, 7, 4}};

// Open-addressing table keyed on up to four sorted vertex ids.
class KeyTable {
 public:
  explicit KeyTable(size_t expected) {
    size_t cap = 16;
    while (cap < 2 * expected + 16) cap <<= 1;
    mask_ = cap - 1;
    keys_.assign(4 * cap, -1);
    vals_.assign(cap, -1);
  }
  // returns existing value or inserts `val` and returns -1
  int find_or_insert(const int* k, int val) {
    size_t h = hash(k) & mask_;
    for (;;) {
      int* s = &keys_[4 * h];
      if (vals_[h] < 0) {
        std::memcpy(s, k, 4 * sizeof(int));
        vals_[h] = val;
        return -1;
      }
      if (s[0] == k[0] && s[1] == k[1] && s[2] == k[2] && s[3] == k[3]) return vals_[h];
      h = (h + 1) & mask_;
    }
  }
  bool contains(const int* k) const {
    size_t h = hash(k) & mask_;
    for (;;) {
      const int* s = &keys_[4 * h];
      if (vals_[h] < 0) return false;
      if (s[0] == k[0] && s[1] == k[1] && s[2] == k[2] && s[3] == k[3]) return true;
      h = (h + 1) & mask_;
    }
  }

 private:
  static size_t hash(const int* k) {
    uint64_t h = 0x9E3779B97F4A7C15ull;
    for (int i = 0; i < 4; ++i) {
      h ^= static_cast<uint32_t>(k[i]) + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
      h *= 0xBF58476D1CE4E5B9ull;
    }
    return static_cast<size_t>(h ^ (h >> 31));
  }
  size_t mask_;
  std::vector<int> keys_;
  std::vector<int> vals_;
};

// sorted distinct vertex key; returns the number of distinct ids
int make_key(int a, int b, int c, int d, int* k) {
  int t[4] = {a, b, c, d};
  std::sort(t, t + 4);
  int n = 0;
  for (int i = 0; i < 4; ++i)
    if (n == 0 || t[i] != k[n - 1]) k[n++] = t[i];
  for (int i = n; i < 4; ++i) k[i] = -1;
  return n;
}

struct InElem {
  int pid;
  int idx;
};
struct ByPart {
  bool operator()(const InElem* l, const InElem* r) const { return l->pid < r->pid ? true : false; }
};

inline void sub3(double* r, const double* a, const double* b) {
  r[0] = a[0] - b[0];
  r[1] = a[1] - b[1];
  r[2] = a[2] - b[2];
}
inline double dot3(const double* a, const double* b) {
  double s = 0.;
  s += a[0] * b[0];
  s += a[1] * b[1];
  s += a[2] * b[2];
  return s;
}
inline double len3(const double* a) { return std::sqrt(dot3(a, a)); }
inline void cross3(double* a, const double* b, const double* c) {
  a[0] = b[1] * c[2] - b[2] * c[1];
  a[1] = -b[0] * c[2] + b[2] * c[0];
  a[2] = b[0] * c[1] - b[1] * c[0];
}
double tri_area(const double* x1, const double* x2, const double* x3) {
  double v1[3], v2[3];
  sub3(v1, x2, x1);
  sub3(v2, x3, x1);
  double a1 = v1[1] * v2[2] - v1[2] * v2[1];
  double a2 = -v1[0] * v2[2] + v1[2] * v2[0];
  double a3 = v1[0] * v2[1] - v1[1] * v2[0];
  return 0.5 * std::sqrt(a1 * a1 + a2 * a2 + a3 * a3);
}
double tet_vol(const double* x1, const double* x2, const double* x3, const double* x4) {
  double v1[3], d1[3], d2[3], d3[3];
  sub3(d1, x2, x1);
  sub3(d2, x3, x1);
  sub3(d3, x4, x1);
  cross3(v1, d1, d2);
  return 1. / 6. * std::fabs(dot3(v1, d3));
}

}  // namespace

int Mesh::type1_of_rid(int rid) const {
  for (const auto& r : regions)
    if (r.rid == rid) return r.type1;
  return -1;
}

namespace {
// S355_VERBOSE=1 prints the time of each set-up stage to stderr.
struct StageClock {
  bool on = std::getenv("S355_VERBOSE") != nullptr;
  std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
  void mark(const char* what) {
    if (!on) return;
    auto n = std::chrono::steady_clock::now();
    std::fprintf(stderr, "[s355 ingest] %-22s %8.3f s\n", what, std::chrono::duration<double>(n - t).count());
    t = n;
  }
};
}  // namespace

Mesh ingest_mesh(const s355_mesh_desc& d) {
  StageClock clk;
  if (d.n_vert <= 0 || d.n_elem <= 0 || !d.xyz || !d.elem_type || !d.elem_tag || !d.elem_vptr || !d.elem_vidx)
    throw std::runtime_error("mesh: empty or missing arrays");
  Mesh m;
  m.n_elem = d.n_elem;
  m.regions.assign(d.regions, d.regions + d.n_region);
  // ---- element order
  {
    std::vector<InElem> el(d.n_elem);
    std::vector<InElem*> ptr(d.n_elem);
    for (int i = 0; i < d.n_elem; ++i) {
      el[i] = {0, i};
      ptr[i] = &el[i];
    }
    std::sort(ptr.begin(), ptr.end(), ByPart());
    m.perm.resize(d.n_elem);
    for (int i = 0; i < d.n_elem; ++i) m.perm[i] = ptr[i]->idx;
  }
  clk.mark("element order");
  // ---- vertex compaction (ascending global id)
  std::vector<int> local(d.n_vert, -1);
  for (int e = 0; e < d.n_elem; ++e) {
    int t = d.elem_type[e];
    if (t < 2 || t > 7) throw std::runtime_error("mesh: unknown element type " + std::to_string(t));
    static const int nn[8] = {0, 0, 3, 4, 4, 8, 6, 5};
    if (d.elem_vptr[e + 1] - d.elem_vptr[e] != nn[t]) throw std::runtime_error("mesh: element vertex count");
    for (int k = d.elem_vptr[e]; k < d.elem_vptr[e + 1]; ++k) {
      int v = d.elem_vidx[k];
      if (v < 0 || v >= d.n_vert) throw std::runtime_error("mesh: vertex id out of range");
      local[v] = 0;
    }
  }
  int nv = 0;
  for (int v = 0; v < d.n_vert; ++v)
    if (local[v] == 0) local[v] = nv++;
  m.n_vert = nv;
  m.vert.resize(3 * static_cast<size_t>(nv));
  for (int v = 0; v < d.n_vert; ++v)
    if (local[v] >= 0)
      for (int k = 0; k < 3; ++k) m.vert[3 * static_cast<size_t>(local[v]) + k] = d.xyz[3 * static_cast<size_t>(v) + k];
  clk.mark("vertex compaction");
  // ---- cells and boundary records
  std::vector<int> cv;     // 8 per cell
  std::vector<int> bv;     // 4 per boundary record
  for (int e : m.perm) {
    int t = d.elem_type[e];
    int v[8];
    const int* src = d.elem_vidx + d.elem_vptr[e];
    int n = d.elem_vptr[e + 1] - d.elem_vptr[e];
    for (int k = 0; k < n; ++k) v[k] = local[src[k]];
    if (t == 2 || t == 3) {
      m.b_rid.push_back(d.elem_tag[e]);
      int q[4] = {v[0], v[1], v[2], t == 2 ? v[2] : v[3]};
      bv.insert(bv.end(), q, q + 4);
    } else {
      int h[8];
      switch (t) {
        case 4: h[0] = v[0]; h[1] = v[1]; h[2] = v[2]; h[3] = v[2]; h[4] = h[5] = h[6] = h[7] = v[3]; break;
        case 5: for (int k = 0; k < 8; ++k) h[k] = v[k]; break;
        case 6:
          h[0] = v[0]; h[1] = v[1]; h[2] = v[2]; h[3] = v[2];
          h[4] = v[3]; h[5] = v[4]; h[6] = v[5]; h[7] = v[5];
          break;
        default: for (int k = 0; k < 4; ++k) h[k] = v[k]; h[4] = h[5] = h[6] = h[7] = v[4]; break;
      }
      m.c_rid.push_back(d.elem_tag[e]);
      cv.insert(cv.end(), h, h + 8);
    }
  }
  m.n_cell = static_cast<int>(m.c_rid.size());
  m.n_bnd = static_cast<int>(m.b_rid.size());
  if (m.n_cell == 0) throw std::runtime_error("mesh: no volume elements");
  for (int r : m.c_rid)
    if (m.type1_of_rid(r) != 5) throw std::runtime_error("mesh: cell region " + std::to_string(r) + " is not a volume");
  for (int r : m.b_rid) {
    int t1 = m.type1_of_rid(r);
    if (t1 < 1 || t1 > 4) throw std::runtime_error("mesh: boundary region " + std::to_string(r) + " unsupported");
  }
  clk.mark("cells/boundary");
  const int Nc = m.n_cell, Nb = m.n_bnd, K = kMaxFace;
  // ---- faces
  KeyTable table(static_cast<size_t>(Nb) + 4 * static_cast<size_t>(Nc));
  m.c_nface.assign(Nc, 0);
  m.c_face.assign(static_cast<size_t>(Nc) * K, -1);
  m.c_cell.assign(static_cast<size_t>(Nc) * K, -1);
  m.c_poly.assign(Nc, 0);
  m.f_c1.reserve(Nb + 3 * static_cast<size_t>(Nc));
  auto new_face = [&](const int* q, int c1, int c2, int b) {
    m.f_v.insert(m.f_v.end(), q, q + 4);
    m.f_c1.push_back(c1);
    m.f_c2.push_back(c2);
    m.f_bnd.push_back(b);
    return static_cast<int>(m.f_c1.size()) - 1;
  };
  for (int i = 0; i < Nb; ++i) {
    int key[4];
    make_key(bv[4 * i], bv[4 * i + 1], bv[4 * i + 2], bv[4 * i + 3], key);
    int f = new_face(&bv[4 * static_cast<size_t>(i)], -9999999, kVoidBoundary, i);
    if (table.find_or_insert(key, f) >= 0) throw std::runtime_error("mesh: duplicate boundary element");
  }
  clk.mark("boundary faces");
  // triangles exposed by cells, for the split-quad extension
  bool any_tri = false, any_quad = false;
  for (int i = 0; i < Nc && !(any_tri && any_quad); ++i)
    for (const auto& p : kPattern) {
      const int* h = &cv[8 * static_cast<size_t>(i)];
      int key[4];
      int n = make_key(h[p[0]], h[p[1]], h[p[2]], h[p[3]], key);
      any_tri |= n == 3;
      any_quad |= n == 4;
    }
  KeyTable tris(any_tri && any_quad ? 4 * static_cast<size_t>(Nc) : 1);
  if (any_tri && any_quad)
    for (int i = 0; i < Nc; ++i)
      for (const auto& p : kPattern) {
        const int* h = &cv[8 * static_cast<size_t>(i)];
        int key[4];
        if (make_key(h[p[0]], h[p[1]], h[p[2]], h[p[3]], key) == 3) tris.find_or_insert(key, 1);
      }
  clk.mark("triangle table");
  auto has_tri = [&](int a, int b, int c) {
    int key[4];
    make_key(a, b, c, c, key);
    return tris.contains(key);
  };
  auto attach = [&](int ic, int n1, int n2, int n3, int n4) {
    int irepeat = 0;
    if (n1 == n2 || n1 == n3 || n1 == n4) irepeat++;
    if (n2 == n3 || n2 == n4) irepeat++;
    if (n3 == n4) irepeat++;
    if (irepeat > 1) return;
    int key[4];
    make_key(n1, n2, n3, n4, key);
    int nf = static_cast<int>(m.f_c1.size());
    int fid = table.find_or_insert(key, nf);
    if (m.c_nface[ic] >= K) throw std::runtime_error("mesh: cell with more than 12 faces");
    if (fid < 0) {
      int q[4] = {n1, n2, n3, n4};
      new_face(q, ic, -20, -10);
      fid = nf;
    } else if (m.f_bnd[fid] >= 0) {
      m.f_c1[fid] = ic;
    } else {
      m.f_c2[fid] = ic;
    }
    m.c_face[static_cast<size_t>(ic) * K + m.c_nface[ic]++] = fid;
  };
  for (int i = 0; i < Nc; ++i) {
    const int* h = &cv[8 * static_cast<size_t>(i)];
    for (const auto& p : kPattern) {
      int n1 = h[p[0]], n2 = h[p[1]], n3 = h[p[2]], n4 = h[p[3]];
      bool quad = n1 != n2 && n1 != n3 && n1 != n4 && n2 != n3 && n2 != n4 && n3 != n4;
      if (quad && any_tri) {
        if (has_tri(n1, n2, n3) && has_tri(n1, n3, n4)) {
          m.c_poly[i] = 1;
          attach(i, n1, n2, n3, n3);
          attach(i, n1, n3, n4, n4);
          continue;
        }
        if (has_tri(n1, n2, n4) && has_tri(n2, n3, n4)) {
          m.c_poly[i] = 1;
          attach(i, n1, n2, n4, n4);
          attach(i, n2, n3, n4, n4);
          continue;
        }
      }
      attach(i, n1, n2, n3, n4);
    }
  }
  clk.mark("interior faces");
  const int Nf = static_cast<int>(m.f_c1.size());
  m.n_face = Nf;
  for (int f = 0; f < Nf; ++f) {
    if (m.f_c2[f] == kVoidBoundary || m.f_c2[f] >= 0) continue;
    throw std::runtime_error("mesh: unmatched face " + std::to_string(f) + " (cell " + std::to_string(m.f_c1[f]) +
                             "); the mesh is not conforming");
  }
  for (int b = 0; b < Nb; ++b)
    if (m.f_c1[b] < 0) throw std::runtime_error("mesh: boundary element " + std::to_string(b) + " has no cell");
  clk.mark("face check");
  // ---- cell metrics
  const double* X = m.vert.data();
  auto P = [&](int v) { return X + 3 * static_cast<size_t>(v); };
  m.c_vol.assign(Nc, 0.0);
  m.c_x.assign(3 * static_cast<size_t>(Nc), 0.0);
  for (int i = 0; i < Nc; ++i) {
    const int* h = &cv[8 * static_cast<size_t>(i)];
    double* xc = &m.c_x[3 * static_cast<size_t>(i)];
    for (int j = 0; j < 3; ++j)
      xc[j] = 1. / 8. * (P(h[0])[j] + P(h[1])[j] + P(h[2])[j] + P(h[3])[j] + P(h[4])[j] + P(h[5])[j] +
                         P(h[6])[j] + P(h[7])[j]);
    if (m.c_poly[i]) continue;
    const double *x1 = P(h[0]), *x2 = P(h[1]), *x3 = P(h[2]), *x4 = P(h[3]), *x5 = P(h[4]), *x6 = P(h[5]),
                 *x7 = P(h[6]), *x8 = P(h[7]);
    double V1 = tet_vol(x1, x2, x4, x6), V2 = tet_vol(x1, x5, x4, x6), V3 = tet_vol(x4, x5, x8, x6);
    double V4 = tet_vol(x2, x3, x4, x7), V5 = tet_vol(x2, x6, x4, x7), V6 = tet_vol(x8, x4, x6, x7);
    m.c_vol[i] = V1 + V2 + V3 + V4 + V5 + V6;
    double c[6][3];
    for (int j = 0; j < 3; ++j) {
      c[0][j] = 0.25 * (x1[j] + x2[j] + x4[j] + x6[j]);
      c[1][j] = 0.25 * (x1[j] + x5[j] + x4[j] + x6[j]);
      c[2][j] = 0.25 * (x4[j] + x5[j] + x8[j] + x6[j]);
      c[3][j] = 0.25 * (x2[j] + x3[j] + x4[j] + x7[j]);
      c[4][j] = 0.25 * (x2[j] + x6[j] + x4[j] + x7[j]);
      c[5][j] = 0.25 * (x8[j] + x4[j] + x6[j] + x7[j]);
    }
    double vs = V1 + V2 + V3 + V4 + V5 + V6;
    for (int j = 0; j < 3; ++j)
      xc[j] = 1. / vs * (V1 * c[0][j] + V2 * c[1][j] + V3 * c[2][j] + V4 * c[3][j] + V5 * c[4][j] + V6 * c[5][j]);
  }
  for (int i = 0; i < Nc; ++i) {
    if (!m.c_poly[i]) continue;
    double* xc = &m.c_x[3 * static_cast<size_t>(i)];
    double xp[3] = {xc[0], xc[1], xc[2]};
    double vs = 0., acc[3] = {0., 0., 0.};
    for (int j = 0; j < m.c_nface[i]; ++j) {
      const int* q = &m.f_v[4 * static_cast<size_t>(m.c_face[static_cast<size_t>(i) * K + j])];
      const int tri[2][3] = {{q[0], q[1], q[3]}, {q[1], q[2], q[3]}};
      for (const auto& t : tri) {
        double v = tet_vol(xp, P(t[0]), P(t[1]), P(t[2]));
        vs += v;
        for (int k = 0; k < 3; ++k) acc[k] += v * 0.25 * (xp[k] + P(t[0])[k] + P(t[1])[k] + P(t[2])[k]);
      }
    }
    m.c_vol[i] = vs;
    for (int k = 0; k < 3; ++k) xc[k] = acc[k] / vs;
  }
  clk.mark("cell metrics");
  for (int i = 0; i < Nc; ++i)
    for (int j = 0; j < m.c_nface[i]; ++j) {
      int f = m.c_face[static_cast<size_t>(i) * K + j];
      int nc = m.f_c1[f] == i ? m.f_c2[f] : m.f_c1[f];
      m.c_cell[static_cast<size_t>(i) * K + j] = nc;
    }
  clk.mark("neighbours");
  // ---- face metrics
  m.f_x.assign(3 * static_cast<size_t>(Nf), 0.0);
  m.f_n.assign(3 * static_cast<size_t>(Nf), 0.0);
  m.f_xpac.assign(3 * static_cast<size_t>(Nf), 0.0);
  m.f_xnac.assign(3 * static_cast<size_t>(Nf), 0.0);
  m.f_area.assign(Nf, 0.0);
  m.f_lambda.assign(Nf, 0.0);
  m.f_rlencos.assign(Nf, 0.0);
  int warped = 0;
  for (int f = 0; f < Nf; ++f) {
    const int* q = &m.f_v[4 * static_cast<size_t>(f)];
    const double *p1 = P(q[0]), *p2 = P(q[1]), *p3 = P(q[2]), *p4 = P(q[3]);
    double a1 = tri_area(p1, p2, p4), a2 = tri_area(p2, p3, p4);
    double* fx = &m.f_x[3 * static_cast<size_t>(f)];
    double* fn = &m.f_n[3 * static_cast<size_t>(f)];
    double xc1[3], xc2[3];
    for (int j = 0; j < 3; ++j) {
      xc1[j] = 1. / 3 * (p1[j] + p2[j] + p4[j]);
      xc2[j] = 1. / 3 * (p2[j] + p3[j] + p4[j]);
    }
    for (int j = 0; j < 3; ++j) fx[j] = (a1 * xc1[j] + a2 * xc2[j]) / (a1 + a2);
    double d1[3], d2[3], dx[3];
    sub3(d1, p1, p3);
    sub3(d2, p2, p4);
    cross3(fn, d1, d2);
    for (int j = 0; j < 3; ++j) fn[j] = 0.5 * fn[j];
    const double* x1 = &m.c_x[3 * static_cast<size_t>(m.f_c1[f])];
    sub3(dx, fx, x1);
    if (dot3(fn, dx) < 0.)
      for (int j = 0; j < 3; ++j) fn[j] = -fn[j];
    double area = len3(fn);
    m.f_area[f] = area;
    if (std::fabs((a1 + a2 - area) / area) > 1.e-4) warped++;
    int c2 = m.f_c2[f];
    if (c2 >= 0) {
      const double* x2 = &m.c_x[3 * static_cast<size_t>(c2)];
      sub3(d1, fx, x1);
      sub3(d2, fx, x2);
      double r1 = len3(d1), r2 = len3(d2);
      m.f_lambda[f] = r2 / (r1 + r2);
      sub3(dx, x2, x1);
      m.f_rlencos[f] = area / std::fabs(dot3(dx, fn) / area);
      double ss;
      sub3(dx, fx, x1);
      ss = dot3(dx, fn) / area;
      for (int j = 0; j < 3; ++j) dx[j] = ss * fn[j] / area;
      sub3(&m.f_xpac[3 * static_cast<size_t>(f)], fx, dx);
      sub3(dx, fx, x2);
      ss = dot3(dx, fn) / area;
      for (int j = 0; j < 3; ++j) dx[j] = ss * fn[j] / area;
      sub3(&m.f_xnac[3 * static_cast<size_t>(f)], fx, dx);
    } else {
      m.f_lambda[f] = 1.;
      sub3(dx, fx, x1);
      m.f_rlencos[f] = area / (dot3(dx, fn) / area);
    }
  }
  clk.mark("face metrics");
  m.n_warped = warped;
  m.b_dist.assign(Nb, 0.0);
  for (int b = 0; b < Nb; ++b) {
    double dx[3];
    sub3(dx, &m.f_x[3 * static_cast<size_t>(b)], &m.c_x[3 * static_cast<size_t>(m.f_c1[b])]);
    m.b_dist[b] = dot3(dx, &m.f_n[3 * static_cast<size_t>(b)]) / m.f_area[b];
    if (m.b_dist[b] < 0.) throw std::runtime_error("mesh: wall distance is negative at boundary " + std::to_string(b));
  }
  clk.mark("boundary distance");
  // ---- CSR pattern
  m.rowptr.assign(Nc + 1, 0);
  m.diag.assign(Nc, -1);
  m.slot.assign(static_cast<size_t>(Nc) * K, -1);
  m.col.reserve(static_cast<size_t>(Nc) + 2 * static_cast<size_t>(Nf - Nb));
  for (int i = 0; i < Nc; ++i) {
    int cols[kMaxFace + 1];
    int nc = 0;
    cols[nc++] = i;
    for (int j = 0; j < m.c_nface[i]; ++j) {
      int nb = m.c_cell[static_cast<size_t>(i) * K + j];
      if (nb >= 0) cols[nc++] = nb;
    }
    std::sort(cols, cols + nc);
    nc = static_cast<int>(std::unique(cols, cols + nc) - cols);
    int base = static_cast<int>(m.col.size());
    for (int k = 0; k < nc; ++k) {
      m.col.push_back(cols[k]);
      if (cols[k] == i) m.diag[i] = base + k;
    }
    for (int j = 0; j < m.c_nface[i]; ++j) {
      int nb = m.c_cell[static_cast<size_t>(i) * K + j];
      if (nb < 0) continue;
      m.slot[static_cast<size_t>(i) * K + j] = base + static_cast<int>(std::lower_bound(cols, cols + nc, nb) - cols);
    }
    m.rowptr[i + 1] = static_cast<int>(m.col.size());
  }
  clk.mark("CSR");
  // ---- compact cell->face lists
  m.cf_ptr.assign(Nc + 1, 0);
  for (int i = 0; i < Nc; ++i) m.cf_ptr[i + 1] = m.cf_ptr[i] + m.c_nface[i];
  m.cf_pat.resize(m.cf_ptr[Nc]);
  m.cf_slot.resize(m.cf_ptr[Nc]);
  m.cf_sorted.resize(m.cf_ptr[Nc]);
  for (int i = 0; i < Nc; ++i) {
    int b = m.cf_ptr[i];
    for (int j = 0; j < m.c_nface[i]; ++j) {
      m.cf_pat[b + j] = m.c_face[static_cast<size_t>(i) * K + j];
      m.cf_slot[b + j] = m.slot[static_cast<size_t>(i) * K + j];
      m.cf_sorted[b + j] = m.cf_pat[b + j];
    }
    std::sort(m.cf_sorted.begin() + b, m.cf_sorted.begin() + b + m.c_nface[i]);
  }
  clk.mark("cell-face lists");
  return m;
}

}  // namespace s355
