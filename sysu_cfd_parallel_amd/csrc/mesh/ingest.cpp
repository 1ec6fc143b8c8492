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

#include "parallel.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <stdexcept>

namespace s355 {
namespace {

const int kPattern[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 1, 5, 4}, {1, 2, 6, 5}, {3, 2, 6, 7}, {0, 3, 7, 4}};

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

struct TriKey {
  int a, b, c;
};

struct HalfFace {
  uint64_t k0, k1;  // sorted distinct vertex ids, -1 padded
  int seq;          // discovery sequence; boundary records are negative
  int pad;
};

// Chunked parallel sort: sort T chunks concurrently, then merge pairs of runs
// in parallel rounds.
template <class Vec, class Cmp>
void parallel_sort(Vec& v, Cmp cmp) {
  const size_t n = v.size();
  int nt = host_threads();
  if (n < (1u << 16) || nt < 2) {
    std::sort(v.begin(), v.end(), cmp);
    return;
  }
  int runs = 1;
  while (runs < nt) runs <<= 1;
  std::vector<size_t> cut(runs + 1);
  for (int r = 0; r <= runs; ++r) cut[r] = n * r / runs;
  parallel_for(runs, [&](long r) { std::sort(v.begin() + cut[r], v.begin() + cut[r + 1], cmp); }, 1);
  Vec tmp(n);
  while (runs > 1) {
    parallel_for(runs / 2, [&](long q) {
      const size_t a = cut[2 * q], b = cut[2 * q + 1], c = cut[2 * q + 2];
      std::merge(v.begin() + a, v.begin() + b, v.begin() + b, v.begin() + c, tmp.begin() + a, cmp);
    }, 1);
    v.swap(tmp);
    for (int r = 0; r <= runs / 2; ++r) cut[r] = cut[2 * r];
    runs /= 2;
  }
}
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

Mesh ingest_mesh(const s355_mesh_desc& d, bool metrics) {
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
  hvec<int>& cv = m.c_vert;  // 8 per cell
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
      m.c_elem.push_back(e);
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
  // Half-faces in discovery order (cells in order, each in pattern order,
  // split quads expanded, degenerate faces skipped), boundary records first.
  // Sorting (vertex key, sequence) puts the two sides of every face next to
  // each other; an interior face takes the rank of its first (creating)
  // half-face, which is exactly the reference's sequential numbering.
  m.c_poly.resize(Nc);
  parallel_for(Nc, [&](long i) { m.c_poly[i] = 0; });
  // triangles exposed by cells, for the split-quad extension
  std::vector<uint8_t> tri_vertex(m.n_vert, 0);
  std::vector<TriKey> tris;
  {
    std::vector<int> tcount(Nc + 1, 0);
    parallel_for(Nc, [&](long i) {
      const int* h = &cv[8 * static_cast<size_t>(i)];
      int c = 0;
      for (const auto& p : kPattern) {
        int key[4];
        if (make_key(h[p[0]], h[p[1]], h[p[2]], h[p[3]], key) == 3) ++c;
      }
      tcount[i + 1] = c;
    });
    for (int i = 0; i < Nc; ++i) tcount[i + 1] += tcount[i];
    bool any_quad = false;
    for (int i = 0; i < Nc && !any_quad; ++i) {
      const int* h = &cv[8 * static_cast<size_t>(i)];
      for (const auto& p : kPattern) {
        int key[4];
        any_quad |= make_key(h[p[0]], h[p[1]], h[p[2]], h[p[3]], key) == 4;
      }
    }
    if (tcount[Nc] > 0 && any_quad) {
      tris.resize(tcount[Nc]);
      parallel_for(Nc, [&](long i) {
        const int* h = &cv[8 * static_cast<size_t>(i)];
        int o = tcount[i];
        for (const auto& p : kPattern) {
          int key[4];
          if (make_key(h[p[0]], h[p[1]], h[p[2]], h[p[3]], key) == 3) tris[o++] = TriKey{key[0], key[1], key[2]};
        }
      });
      parallel_sort(tris, [](const TriKey& a, const TriKey& b) {
        return a.a != b.a ? a.a < b.a : (a.b != b.b ? a.b < b.b : a.c < b.c);
      });
      tris.erase(std::unique(tris.begin(), tris.end(),
                             [](const TriKey& a, const TriKey& b) { return a.a == b.a && a.b == b.b && a.c == b.c; }),
                 tris.end());
      for (const TriKey& t : tris) tri_vertex[t.a] = tri_vertex[t.b] = tri_vertex[t.c] = 1;
    }
  }
  clk.mark("triangle set");
  auto has_tri = [&](int a, int b, int c) {
    int key[4];
    make_key(a, b, c, c, key);
    TriKey t{key[0], key[1], key[2]};
    return std::binary_search(tris.begin(), tris.end(), t, [](const TriKey& x, const TriKey& y) {
      return x.a != y.a ? x.a < y.a : (x.b != y.b ? x.b < y.b : x.c < y.c);
    });
  };
  // expand cell i's faces; emit(n1, n2, n3, n4) per non-degenerate face
  auto expand = [&](long i, auto&& emit) {
    const int* h = &cv[8 * static_cast<size_t>(i)];
    auto put = [&](int n1, int n2, int n3, int n4) {
      int irepeat = 0;
      if (n1 == n2 || n1 == n3 || n1 == n4) irepeat++;
      if (n2 == n3 || n2 == n4) irepeat++;
      if (n3 == n4) irepeat++;
      if (irepeat > 1) return;
      emit(n1, n2, n3, n4);
    };
    for (const auto& p : kPattern) {
      int n1 = h[p[0]], n2 = h[p[1]], n3 = h[p[2]], n4 = h[p[3]];
      bool quad = n1 != n2 && n1 != n3 && n1 != n4 && n2 != n3 && n2 != n4 && n3 != n4;
      if (quad && !tris.empty() && tri_vertex[n1] && tri_vertex[n2] && tri_vertex[n3] && tri_vertex[n4]) {
        if (has_tri(n1, n2, n3) && has_tri(n1, n3, n4)) {
          m.c_poly[i] = 1;
          put(n1, n2, n3, n3);
          put(n1, n3, n4, n4);
          continue;
        }
        if (has_tri(n1, n2, n4) && has_tri(n2, n3, n4)) {
          m.c_poly[i] = 1;
          put(n1, n2, n4, n4);
          put(n2, n3, n4, n4);
          continue;
        }
      }
      put(n1, n2, n3, n4);
    }
  };
  m.c_nface.resize(Nc);
  m.cf_ptr.resize(Nc + 1);
  m.cf_ptr[0] = 0;
  parallel_for(Nc, [&](long i) {
    int c = 0;
    expand(i, [&](int, int, int, int) { ++c; });
    m.c_nface[i] = c;
  });
  {
    // cell-face entries (and so half-face records) are indexed by 32 bits
    long long total = 0;
    for (int i = 0; i < Nc; ++i) {
      if (m.c_nface[i] > K) throw std::runtime_error("mesh: cell with more than 12 faces");
      total += m.c_nface[i];
    }
    if (total > std::numeric_limits<int>::max())
      throw std::runtime_error("mesh: more cell faces than 32-bit indices address");
  }
  for (int i = 0; i < Nc; ++i) m.cf_ptr[i + 1] = m.cf_ptr[i] + m.c_nface[i];
  const long H = m.cf_ptr[Nc];
  hvec<int> hv(4 * static_cast<size_t>(H));   // vertex order of each half-face
  hvec<int> hcell(H);                          // owning cell
  hvec<HalfFace> rec(static_cast<size_t>(Nb) + H);
  auto fill_key = [](HalfFace& r, int a, int b, int c, int d, int seq) {
    int key[4];
    make_key(a, b, c, d, key);
    r.k0 = (static_cast<uint64_t>(static_cast<uint32_t>(key[0])) << 32) | static_cast<uint32_t>(key[1]);
    r.k1 = (static_cast<uint64_t>(static_cast<uint32_t>(key[2])) << 32) | static_cast<uint32_t>(key[3]);
    r.seq = seq;
  };
  parallel_for(Nb, [&](long b) {
    const int* q = &bv[4 * static_cast<size_t>(b)];
    fill_key(rec[b], q[0], q[1], q[2], q[3], static_cast<int>(b - Nb));
  });
  parallel_for(Nc, [&](long i) {
    long e = m.cf_ptr[i];
    expand(i, [&](int n1, int n2, int n3, int n4) {
      int* o = &hv[4 * static_cast<size_t>(e)];
      o[0] = n1;
      o[1] = n2;
      o[2] = n3;
      o[3] = n4;
      hcell[e] = static_cast<int>(i);
      fill_key(rec[Nb + e], n1, n2, n3, n4, static_cast<int>(e));
      ++e;
    });
  });
  clk.mark("half-faces");
  parallel_sort(rec, [](const HalfFace& a, const HalfFace& b) {
    return a.k0 != b.k0 ? a.k0 < b.k0 : (a.k1 != b.k1 ? a.k1 < b.k1 : a.seq < b.seq);
  });
  clk.mark("half-face sort");
  // pair the sides: partner[e1] = e2 for interior faces created by e1; bface[e] = b for boundary
  hvec<int> partner(H), face_of(H);
  hvec<uint8_t> creator(H);
  parallel_for(H, [&](long e) { creator[e] = 0; });
  std::vector<int> bcell(Nb, -1);
  {
    const size_t M = rec.size();
    size_t s0 = 0;
    while (s0 < M) {
      size_t t = s0 + 1;
      while (t < M && rec[t].k0 == rec[s0].k0 && rec[t].k1 == rec[s0].k1) ++t;
      const HalfFace& a = rec[s0];
      if (t - s0 > 2) throw std::runtime_error("mesh: a face is shared by more than two elements");
      if (t - s0 == 1) {
        if (a.seq < 0)
          throw std::runtime_error("mesh: boundary element " + std::to_string(a.seq + Nb) + " has no cell");
        throw std::runtime_error("mesh: unmatched face of cell " + std::to_string(hcell[a.seq]) +
                                 "; the mesh is not conforming");
      }
      const HalfFace& b = rec[s0 + 1];
      if (a.seq < 0) {
        if (b.seq < 0) throw std::runtime_error("mesh: duplicate boundary element");
        face_of[b.seq] = a.seq + Nb;
        bcell[a.seq + Nb] = hcell[b.seq];
      } else {
        creator[a.seq] = 1;
        partner[a.seq] = b.seq;
      }
      s0 = t;
    }
  }
  hvec<HalfFace>().swap(rec);
  int Nf = Nb;
  for (long e = 0; e < H; ++e)
    if (creator[e]) face_of[e] = Nf++;
  m.f_c1.resize(Nf);
  m.f_c2.resize(Nf);
  m.f_bnd.resize(Nf);
  m.f_v.resize(4 * static_cast<size_t>(Nf));
  for (int b = 0; b < Nb; ++b) {
    m.f_c1[b] = bcell[b];
    m.f_c2[b] = kVoidBoundary;
    m.f_bnd[b] = b;
    for (int k = 0; k < 4; ++k) m.f_v[4 * static_cast<size_t>(b) + k] = bv[4 * static_cast<size_t>(b) + k];
  }
  parallel_for(H, [&](long e) {
    if (!creator[e]) return;
    const int f = face_of[e], e2 = partner[e];
    face_of[e2] = f;
    m.f_c1[f] = hcell[e];
    m.f_c2[f] = hcell[e2];
    m.f_bnd[f] = -10;
    for (int k = 0; k < 4; ++k) m.f_v[4 * static_cast<size_t>(f) + k] = hv[4 * static_cast<size_t>(e) + k];
  });
  m.c_face.resize(static_cast<size_t>(Nc) * K);
  m.cf_pat.resize(H);
  parallel_for(Nc, [&](long i) {
    for (int j = 0; j < K; ++j) {
      const int f = j < m.c_nface[i] ? face_of[m.cf_ptr[i] + j] : -1;
      m.c_face[static_cast<size_t>(i) * K + j] = f;
      if (j < m.c_nface[i]) m.cf_pat[m.cf_ptr[i] + j] = f;
    }
  });
  m.n_face = Nf;
  clk.mark("face numbering");
  // ---- cell metrics
  const double* X = m.vert.data();
  auto P = [&](int v) { return X + 3 * static_cast<size_t>(v); };
  if (metrics) {
    m.c_vol.resize(Nc);
    m.c_x.resize(3 * static_cast<size_t>(Nc));
    parallel_for(Nc, [&](long i) {
      const int* h = &cv[8 * static_cast<size_t>(i)];
      double* xc = &m.c_x[3 * static_cast<size_t>(i)];
      for (int j = 0; j < 3; ++j)
        xc[j] = 1. / 8. * (P(h[0])[j] + P(h[1])[j] + P(h[2])[j] + P(h[3])[j] + P(h[4])[j] + P(h[5])[j] +
                           P(h[6])[j] + P(h[7])[j]);
      if (m.c_poly[i]) {
        // tets on the face triangles about the vertex mean
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
        return;
      }
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
    });
  }
  m.c_cell.resize(static_cast<size_t>(Nc) * K);
  parallel_for(Nc, [&](long i) {
    for (int j = 0; j < K; ++j) {
      int nc = -1;
      if (j < m.c_nface[i]) {
        int f = m.c_face[static_cast<size_t>(i) * K + j];
        nc = m.f_c1[f] == i ? m.f_c2[f] : m.f_c1[f];
      }
      m.c_cell[static_cast<size_t>(i) * K + j] = nc;
    }
  });
  clk.mark("cell metrics");
  // ---- face metrics
  if (metrics) {
    m.f_x.resize(3 * static_cast<size_t>(Nf));
    m.f_n.resize(3 * static_cast<size_t>(Nf));
    m.f_xpac.resize(3 * static_cast<size_t>(Nf));
    m.f_xnac.resize(3 * static_cast<size_t>(Nf));
    m.f_area.resize(Nf);
    m.f_lambda.resize(Nf);
    m.f_rlencos.resize(Nf);
    std::atomic<int> warped_sum{0};
    parallel_chunks(Nf, [&](long f0, long f1) {
      int warped = 0;
      for (long f = f0; f < f1; ++f) {
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
          for (int j = 0; j < 3; ++j)
            m.f_xpac[3 * static_cast<size_t>(f) + j] = m.f_xnac[3 * static_cast<size_t>(f) + j] = 0.0;
        }
      }
      warped_sum += warped;
    });
    m.n_warped = warped_sum.load();
    clk.mark("face metrics");
    m.b_dist.resize(Nb);
    for (int b = 0; b < Nb; ++b) {
      double dx[3];
      sub3(dx, &m.f_x[3 * static_cast<size_t>(b)], &m.c_x[3 * static_cast<size_t>(m.f_c1[b])]);
      m.b_dist[b] = dot3(dx, &m.f_n[3 * static_cast<size_t>(b)]) / m.f_area[b];
      if (m.b_dist[b] < 0.)
        throw std::runtime_error("mesh: wall distance is negative at boundary " + std::to_string(b));
    }
  }
  // ---- CSR pattern: count, prefix, fill
  m.rowptr.resize(Nc + 1);
  m.rowptr[0] = 0;
  m.diag.resize(Nc);
  m.slot.resize(static_cast<size_t>(Nc) * K);
  auto row_cols = [&](long i, int* cols) {
    int nc = 0;
    cols[nc++] = static_cast<int>(i);
    for (int j = 0; j < m.c_nface[i]; ++j) {
      int nb = m.c_cell[static_cast<size_t>(i) * K + j];
      if (nb >= 0) cols[nc++] = nb;
    }
    std::sort(cols, cols + nc);
    return static_cast<int>(std::unique(cols, cols + nc) - cols);
  };
  parallel_for(Nc, [&](long i) {
    int cols[kMaxFace + 1];
    m.rowptr[i + 1] = row_cols(i, cols);
  });
  {
    long long nnz = 0;
    for (int i = 0; i < Nc; ++i) nnz += m.rowptr[i + 1];
    if (nnz > std::numeric_limits<int>::max())
      throw std::runtime_error("mesh: more matrix entries than 32-bit CSR indices address");
  }
  for (int i = 0; i < Nc; ++i) m.rowptr[i + 1] += m.rowptr[i];
  m.col.resize(m.rowptr[Nc]);
  parallel_for(Nc, [&](long i) {
    int cols[kMaxFace + 1];
    const int nc = row_cols(i, cols);
    const int base = m.rowptr[i];
    for (int k = 0; k < nc; ++k) {
      m.col[base + k] = cols[k];
      if (cols[k] == i) m.diag[i] = base + k;
    }
    for (int j = 0; j < K; ++j) {
      int nb = m.c_cell[static_cast<size_t>(i) * K + j];
      m.slot[static_cast<size_t>(i) * K + j] =
          nb < 0 ? -1 : base + static_cast<int>(std::lower_bound(cols, cols + nc, nb) - cols);
    }
  });
  clk.mark("CSR");
  // ---- compact cell->face lists (cf_ptr / cf_pat were filled with the faces)
  m.cf_slot.resize(m.cf_ptr[Nc]);
  m.cf_sorted.resize(m.cf_ptr[Nc]);
  parallel_for(Nc, [&](long i) {
    const int b = m.cf_ptr[i];
    for (int j = 0; j < m.c_nface[i]; ++j) {
      m.cf_slot[b + j] = m.slot[static_cast<size_t>(i) * K + j];
      m.cf_sorted[b + j] = m.cf_pat[b + j];
    }
    std::sort(m.cf_sorted.begin() + b, m.cf_sorted.begin() + b + m.c_nface[i]);
  });
  clk.mark("cell-face lists");
  return m;
}

}  // namespace s355
