// ORACLE — test infrastructure only (see oracle.hpp).
//
// Mesh ingest restated from the reference:
//   element order  RootStructure.cpp:797-798 (std::sort on equal partition
//                  keys), vertex compaction RootStructure.cpp:930-968,
//                  1026-1029;
//   cells/bnds     geometry.cpp:105-203 (every element as an 8-node hex);
//   faces          geometry.cpp:222-358 (boundary faces first, then
//                  discovery order, pattern (0123)(4567)(0154)(1265)(3267)(0374));
//   metrics        geometry.cpp:391-752;  check geometry.cpp:754-768;
//   CSR pattern    MatMPIAIJSetPreallocation + MatSetValues rows
//                  (MPIStructure.cpp:29-44, navier_velocity.cpp:365-371).
// Deviations (both documented as relaxations; the reference aborts instead):
//   * the planarity check of geometry.cpp:604-610 counts warped quads
//     (n_warped) rather than calling exit(0);
//   * a hex quad covered by two tetrahedron triangles is split into those two
//     triangles (polyhedral cell); such cells take volume and centroid from a
//     face-triangle decomposition about the vertex mean.  The reference cannot
//     represent these meshes (cells hold at most 6 faces).
// Parity pins: closed forms on uniform cubes (vol, area, lambda, rlencos,
// distance, Xpac/Xnac), tet volume, sum of volumes, closed-cell normal sums,
// face counts of 1x1x1 / 2x1x1 hex and one / two tets, the 40-key sort
// permutation probe, and integer equality with the HIP library's builder.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <set>
#include <stdexcept>

#include "oracle.hpp"

namespace orc {
namespace {

void vminus(double* r, const double* a, const double* b) {
  for (int i = 0; i < 3; ++i) r[i] = a[i] - b[i];
}
double vdot(const double* a, const double* b) {
  double s = 0.;
  for (int i = 0; i < 3; ++i) s += a[i] * b[i];
  return s;
}
double vlen(const double* a) {
  double s = 0.;
  for (int i = 0; i < 3; ++i) s += a[i] * a[i];
  return std::sqrt(s);
}
void vcross(double* a, const double* b, const double* c) {
  a[0] = b[1] * c[2] - b[2] * c[1];
  a[1] = -b[0] * c[2] + b[2] * c[0];
  a[2] = b[0] * c[1] - b[1] * c[0];
}
double tri_area(const double* x1, const double* x2, const double* x3) {
  double v1[3], v2[3];
  vminus(v1, x2, x1);
  vminus(v2, x3, x1);
  double a1 = v1[1] * v2[2] - v1[2] * v2[1];
  double a2 = -v1[0] * v2[2] + v1[2] * v2[0];
  double a3 = v1[0] * v2[1] - v1[1] * v2[0];
  return 0.5 * std::sqrt(a1 * a1 + a2 * a2 + a3 * a3);
}
double tet_vol(const double* x1, const double* x2, const double* x3, const double* x4) {
  double v1[3], d1[3], d2[3], d3[3];
  vminus(d1, x2, x1);
  vminus(d2, x3, x1);
  vminus(d3, x4, x1);
  vcross(v1, d1, d2);
  return 1. / 6. * std::fabs(vdot(v1, d3));
}

struct InElem {
  int pid;
  int idx;
};
struct ByPart {
  bool operator()(const InElem* l, const InElem* r) const { return l->pid < r->pid ? true : false; }
};

const int kPattern[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 1, 5, 4}, {1, 2, 6, 5}, {3, 2, 6, 7}, {0, 3, 7, 4}};

}  // namespace

int Oracle::type1_of(int rid) const {
  auto it = region.find(rid);
  if (it == region.end()) throw std::runtime_error("unknown region id " + std::to_string(rid));
  return it->second.type1;
}

void Oracle::ingest(const s355_mesh_desc& m) {
  if (m.n_vert <= 0 || m.n_elem <= 0) throw std::runtime_error("empty mesh");
  for (int r = 0; r < m.n_region; ++r) region[m.regions[r].rid] = m.regions[r];
  // element order after the single-partition sort (all keys equal)
  std::vector<InElem> el(m.n_elem);
  std::vector<InElem*> ptr(m.n_elem);
  for (int i = 0; i < m.n_elem; ++i) {
    el[i] = {0, i};
    ptr[i] = &el[i];
  }
  std::sort(ptr.begin(), ptr.end(), ByPart());
  perm.resize(m.n_elem);
  for (int i = 0; i < m.n_elem; ++i) perm[i] = ptr[i]->idx;
  // vertex pool: ordered map over the vertices the elements use
  std::map<int, int> pool;
  for (int e : perm)
    for (int k = m.elem_vptr[e]; k < m.elem_vptr[e + 1]; ++k) pool.insert(std::make_pair(m.elem_vidx[k], 0));
  int lid = 0;
  vert.resize(3 * pool.size());
  for (auto& kv : pool) {
    kv.second = lid;
    for (int d = 0; d < 3; ++d) vert[3 * lid + d] = m.xyz[3 * kv.first + d];
    ++lid;
  }
  Nvrt = lid;
  // cells and boundary records in sorted element order
  for (int e : perm) {
    int t = m.elem_type[e];
    int nv = m.elem_vptr[e + 1] - m.elem_vptr[e];
    int v[8];
    for (int k = 0; k < nv && k < 8; ++k) v[k] = pool[m.elem_vidx[m.elem_vptr[e] + k]];
    if (t == 2 || t == 3) {
      Bnd b;
      b.rid = m.elem_tag[e];
      b.v[0] = v[0];
      b.v[1] = v[1];
      b.v[2] = v[2];
      b.v[3] = (t == 2) ? v[2] : v[3];
      bnd.push_back(b);
    } else if (t >= 4 && t <= 7) {
      Cell c;
      c.rid = m.elem_tag[e];
      if (t == 4) {
        for (int p = 0; p < 3; ++p) c.v[p] = v[p];
        c.v[3] = v[2];
        for (int p = 4; p < 8; ++p) c.v[p] = v[3];
      } else if (t == 5) {
        for (int p = 0; p < 8; ++p) c.v[p] = v[p];
      } else if (t == 6) {
        for (int p = 0; p < 3; ++p) c.v[p] = v[p];
        c.v[3] = v[2];
        for (int p = 4; p < 7; ++p) c.v[p] = v[p - 1];
        c.v[7] = v[5];
      } else {
        for (int p = 0; p < 4; ++p) c.v[p] = v[p];
        for (int p = 4; p < 8; ++p) c.v[p] = v[4];
      }
      cell.push_back(c);
    } else {
      throw std::runtime_error("unknown element type " + std::to_string(t));
    }
  }
  Ncel = static_cast<int>(cell.size());
  Nbnd = static_cast<int>(bnd.size());
  for (const auto& c : cell) type1_of(c.rid);
  for (const auto& b : bnd) {
    int t1 = type1_of(b.rid);
    if (t1 < 1 || t1 > 4) throw std::runtime_error("unsupported boundary type " + std::to_string(t1));
  }
}

void Oracle::find_face(int ic, int n1, int n2, int n3, int n4, std::vector<std::vector<int>>& node_face) {
  int irepeat = 0;
  if (n1 == n2 || n1 == n3 || n1 == n4) irepeat++;
  if (n2 == n3 || n2 == n4) irepeat++;
  if (n3 == n4) irepeat++;
  if (irepeat > 1) return;
  int fid = -100;
  for (int f : node_face[n1]) {
    const int* q = face[f].v;
    bool in1 = q[0] == n1 || q[0] == n2 || q[0] == n3 || q[0] == n4;
    bool in2 = q[1] == n1 || q[1] == n2 || q[1] == n3 || q[1] == n4;
    bool in3 = q[2] == n1 || q[2] == n2 || q[2] == n3 || q[2] == n4;
    bool in4 = q[3] == n1 || q[3] == n2 || q[3] == n3 || q[3] == n4;
    if (in1 && in2 && in3 && in4) {
      fid = f;
      break;
    }
  }
  Cell& c = cell[ic];
  if (c.nface >= kMaxFace) throw std::runtime_error("cell with too many faces");
  if (fid < 0) {
    Face f;
    f.v[0] = n1;
    f.v[1] = n2;
    f.v[2] = n3;
    f.v[3] = n4;
    f.c1 = ic;
    f.c2 = kVoidInterface;
    f.bnd = kInnerFace;
    int nf = static_cast<int>(face.size());
    face.push_back(f);
    c.face[c.nface++] = nf;
    for (int k = 0; k < 4; ++k) {
      auto& lst = node_face[face[nf].v[k]];
      if (lst.empty() || lst.back() != nf) lst.push_back(nf);  // ascending, like std::set<int>
    }
  } else {
    if (face[fid].bnd >= 0)
      face[fid].c1 = ic;
    else
      face[fid].c2 = ic;
    c.face[c.nface++] = fid;
  }
}

void Oracle::create_faces() {
  std::vector<std::vector<int>> node_face(Nvrt);
  for (auto& c : cell) c.nface = 0;
  face.clear();
  face.reserve(static_cast<size_t>(Nbnd) + 3 * static_cast<size_t>(Ncel));
  for (int i = 0; i < Nbnd; ++i) {
    bnd[i].face = static_cast<int>(face.size());
    Face f;
    for (int j = 0; j < 4; ++j) f.v[j] = bnd[i].v[j];
    f.bnd = i;
    f.c2 = kVoidBoundary;
    f.c1 = -9999999;
    int nf = static_cast<int>(face.size());
    face.push_back(f);
    for (int j = 0; j < 4; ++j) {
      auto& lst = node_face[f.v[j]];
      if (lst.empty() || lst.back() != nf) lst.push_back(nf);
    }
  }
  // Polyhedral extension: triangles that cells expose, to detect hex quads
  // covered by two triangles of split neighbours.
  std::set<std::array<int, 3>> tris;
  for (const auto& c : cell)
    for (const auto& p : kPattern) {
      int a[4] = {c.v[p[0]], c.v[p[1]], c.v[p[2]], c.v[p[3]]};
      std::set<int> s(a, a + 4);
      if (s.size() == 3) {
        std::array<int, 3> t;
        std::copy(s.begin(), s.end(), t.begin());
        tris.insert(t);
      }
    }
  auto has_tri = [&](int a, int b, int c) {
    std::array<int, 3> t = {a, b, c};
    std::sort(t.begin(), t.end());
    return tris.count(t) > 0;
  };
  for (int i = 0; i < Ncel; ++i) {
    Cell& c = cell[i];
    for (const auto& p : kPattern) {
      int n1 = c.v[p[0]], n2 = c.v[p[1]], n3 = c.v[p[2]], n4 = c.v[p[3]];
      bool quad = n1 != n2 && n1 != n3 && n1 != n4 && n2 != n3 && n2 != n4 && n3 != n4;
      if (quad && !tris.empty()) {
        if (has_tri(n1, n2, n3) && has_tri(n1, n3, n4)) {
          c.poly = true;
          find_face(i, n1, n2, n3, n3, node_face);
          find_face(i, n1, n3, n4, n4, node_face);
          continue;
        }
        if (has_tri(n1, n2, n4) && has_tri(n2, n3, n4)) {
          c.poly = true;
          find_face(i, n1, n2, n4, n4, node_face);
          find_face(i, n2, n3, n4, n4, node_face);
          continue;
        }
      }
      find_face(i, n1, n2, n3, n4, node_face);
    }
  }
  Nfac = static_cast<int>(face.size());
}

void Oracle::cell_face_info() {
  auto X = [&](int v) { return &vert[3 * v]; };
  for (int i = 0; i < Ncel; ++i) {
    Cell& c = cell[i];
    for (int j = 0; j < 3; ++j)
      c.x[j] = 1. / 8. *
               (X(c.v[0])[j] + X(c.v[1])[j] + X(c.v[2])[j] + X(c.v[3])[j] + X(c.v[4])[j] + X(c.v[5])[j] +
                X(c.v[6])[j] + X(c.v[7])[j]);
  }
  for (int i = 0; i < Ncel; ++i) {
    Cell& c = cell[i];
    if (c.poly) continue;  // done below, once the faces carry their vertices
    const double *x1 = X(c.v[0]), *x2 = X(c.v[1]), *x3 = X(c.v[2]), *x4 = X(c.v[3]), *x5 = X(c.v[4]),
                 *x6 = X(c.v[5]), *x7 = X(c.v[6]), *x8 = X(c.v[7]);
    double V1 = tet_vol(x1, x2, x4, x6);
    double V2 = tet_vol(x1, x5, x4, x6);
    double V3 = tet_vol(x4, x5, x8, x6);
    double V4 = tet_vol(x2, x3, x4, x7);
    double V5 = tet_vol(x2, x6, x4, x7);
    double V6 = tet_vol(x8, x4, x6, x7);
    c.vol = V1 + V2 + V3 + V4 + V5 + V6;
    double xc[6][3];
    for (int j = 0; j < 3; ++j) {
      xc[0][j] = 0.25 * (x1[j] + x2[j] + x4[j] + x6[j]);
      xc[1][j] = 0.25 * (x1[j] + x5[j] + x4[j] + x6[j]);
      xc[2][j] = 0.25 * (x4[j] + x5[j] + x8[j] + x6[j]);
      xc[3][j] = 0.25 * (x2[j] + x3[j] + x4[j] + x7[j]);
      xc[4][j] = 0.25 * (x2[j] + x6[j] + x4[j] + x7[j]);
      xc[5][j] = 0.25 * (x8[j] + x4[j] + x6[j] + x7[j]);
    }
    double vs = V1 + V2 + V3 + V4 + V5 + V6;
    for (int j = 0; j < 3; ++j)
      c.x[j] = 1. / vs * (V1 * xc[0][j] + V2 * xc[1][j] + V3 * xc[2][j] + V4 * xc[3][j] + V5 * xc[4][j] +
                          V6 * xc[5][j]);
  }
  // polyhedral cells: tets on face triangles about the vertex mean
  for (int i = 0; i < Ncel; ++i) {
    Cell& c = cell[i];
    if (!c.poly) continue;
    double xp[3] = {c.x[0], c.x[1], c.x[2]};
    double vs = 0., acc[3] = {0., 0., 0.};
    for (int j = 0; j < c.nface; ++j) {
      const int* q = face[c.face[j]].v;
      const int tri[2][3] = {{q[0], q[1], q[3]}, {q[1], q[2], q[3]}};
      for (const auto& t : tri) {
        double v = tet_vol(xp, X(t[0]), X(t[1]), X(t[2]));
        vs += v;
        for (int d = 0; d < 3; ++d) acc[d] += v * 0.25 * (xp[d] + X(t[0])[d] + X(t[1])[d] + X(t[2])[d]);
      }
    }
    c.vol = vs;
    for (int d = 0; d < 3; ++d) c.x[d] = acc[d] / vs;
  }
  // neighbours
  for (int i = 0; i < Ncel; ++i) {
    Cell& c = cell[i];
    for (int j = 0; j < c.nface; ++j) {
      int f = c.face[j];
      int nc = face[f].c1;
      if (nc == i) nc = face[f].c2;
      c.cell[j] = nc;
    }
    for (int j = c.nface; j < kMaxFace; ++j) c.face[j] = c.cell[j] = -1;
  }
  // faces
  n_warped = 0;
  for (int i = 0; i < Nfac; ++i) {
    Face& f = face[i];
    const double *p1 = X(f.v[0]), *p2 = X(f.v[1]), *p3 = X(f.v[2]), *p4 = X(f.v[3]);
    double a1 = tri_area(p1, p2, p4);
    double a2 = tri_area(p2, p3, p4);
    double xc1[3], xc2[3];
    for (int j = 0; j < 3; ++j) {
      xc1[j] = 1. / 3 * (p1[j] + p2[j] + p4[j]);
      xc2[j] = 1. / 3 * (p2[j] + p3[j] + p4[j]);
    }
    for (int j = 0; j < 3; ++j) f.x[j] = (a1 * xc1[j] + a2 * xc2[j]) / (a1 + a2);
    double d1[3], d2[3], dx[3];
    vminus(d1, p1, p3);
    vminus(d2, p2, p4);
    vcross(f.n, d1, d2);
    for (int j = 0; j < 3; ++j) f.n[j] = 0.5 * f.n[j];
    vminus(dx, f.x, cell[f.c1].x);
    if (vdot(f.n, dx) < 0.)
      for (int j = 0; j < 3; ++j) f.n[j] = -f.n[j];
    f.area = vlen(f.n);
    if (std::fabs((a1 + a2 - f.area) / f.area) > 1.e-4) n_warped++;
    if (f.c2 >= 0) {
      vminus(d1, f.x, cell[f.c1].x);
      vminus(d2, f.x, cell[f.c2].x);
      double r1 = vlen(d1), r2 = vlen(d2);
      f.lambda = r2 / (r1 + r2);
    } else {
      f.lambda = 1.;
    }
  }
  for (int i = 0; i < Nbnd; ++i) {
    int jf = bnd[i].face;
    int jc = face[jf].c1;
    double dx[3];
    vminus(dx, face[jf].x, cell[jc].x);
    bnd[i].distance = vdot(dx, face[jf].n) / face[jf].area;
    if (bnd[i].distance < 0.) throw std::runtime_error("wall distance is negative: " + std::to_string(i));
  }
  for (int i = 0; i < Nfac; ++i) {
    Face& f = face[i];
    double dx[3];
    if (f.c2 == kVoidBoundary) {
      vminus(dx, f.x, cell[f.c1].x);
      f.rlencos = f.area / (vdot(dx, f.n) / f.area);
    } else {
      vminus(dx, cell[f.c2].x, cell[f.c1].x);
      f.rlencos = f.area / std::fabs(vdot(dx, f.n) / f.area);
    }
    if (f.c2 >= 0) {
      double ss;
      vminus(dx, f.x, cell[f.c1].x);
      ss = vdot(dx, f.n) / f.area;
      for (int j = 0; j < 3; ++j) dx[j] = ss * f.n[j] / f.area;
      vminus(f.xpac, f.x, dx);
      vminus(dx, f.x, cell[f.c2].x);
      ss = vdot(dx, f.n) / f.area;
      for (int j = 0; j < 3; ++j) dx[j] = ss * f.n[j] / f.area;
      vminus(f.xnac, f.x, dx);
    }
  }
}

void Oracle::check_faces() {
  for (int i = 0; i < Nfac; ++i) {
    int c2 = face[i].c2;
    if (c2 == kVoidBoundary || c2 >= 0) continue;
    char msg[160];
    std::snprintf(msg, sizeof msg, "error in face right hand side: face %d cell1 %d cell2 %d", i, face[i].c1, c2);
    throw std::runtime_error(msg);
  }
  for (int i = 0; i < Nbnd; ++i)
    if (face[bnd[i].face].c1 < 0) throw std::runtime_error("boundary element not attached to a cell");
}

void Oracle::build_csr() {
  csr.rowptr.assign(Ncel + 1, 0);
  csr.col.clear();
  csr.diag.assign(Ncel, -1);
  csr.slot.assign(static_cast<size_t>(Ncel) * kMaxFace, -1);
  for (int i = 0; i < Ncel; ++i) {
    std::vector<int> cols(1, i);
    for (int j = 0; j < cell[i].nface; ++j)
      if (cell[i].cell[j] >= 0) cols.push_back(cell[i].cell[j]);
    std::sort(cols.begin(), cols.end());
    cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
    int base = static_cast<int>(csr.col.size());
    for (size_t k = 0; k < cols.size(); ++k) {
      csr.col.push_back(cols[k]);
      if (cols[k] == i) csr.diag[i] = base + static_cast<int>(k);
    }
    for (int j = 0; j < cell[i].nface; ++j) {
      int nb = cell[i].cell[j];
      if (nb < 0) continue;
      int pos = static_cast<int>(std::lower_bound(cols.begin(), cols.end(), nb) - cols.begin());
      csr.slot[static_cast<size_t>(i) * kMaxFace + j] = base + pos;
    }
    csr.rowptr[i + 1] = static_cast<int>(csr.col.size());
  }
  cell_faces_sorted.assign(static_cast<size_t>(Ncel) * kMaxFace, -1);
  for (int i = 0; i < Ncel; ++i) {
    std::vector<int> fs(cell[i].face, cell[i].face + cell[i].nface);
    std::sort(fs.begin(), fs.end());
    for (size_t k = 0; k < fs.size(); ++k) cell_faces_sorted[static_cast<size_t>(i) * kMaxFace + k] = fs[k];
  }
}

}  // namespace orc
