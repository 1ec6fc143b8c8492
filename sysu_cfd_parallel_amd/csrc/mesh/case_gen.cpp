// Synthetic mesh generators producing an element stream in gmsh order:
// boundary elements (tri/quad) first, then volume elements.  Meshes are
// lattice blocks of hexahedra with seeded node jitter; a hex may be split into
// the six tetrahedra of the Kuhn (Freudenthal) subdivision, which is
// conforming between split neighbours because every face is cut along the
// diagonal through the cell's (0,0,0) or (1,1,1) corner.
#include <array>
#include <cmath>
#include <stdexcept>
#include <vector>

#include "element_stream.hpp"

namespace s355 {
namespace {

// corner bit pattern c = a | b<<1 | c<<2 (a along x, b along y, c along z)
// gmsh hex order: (000)(100)(110)(010)(001)(101)(111)(011)
constexpr int kHexCorner[8] = {0, 1, 3, 2, 4, 5, 7, 6};

struct Lattice {
  int nx, ny, nz;
  std::vector<double> xs, ys, zs;  // node coordinates along each axis
  std::vector<uint8_t> active;     // per cell
  std::vector<uint8_t> split;      // per cell
  long cid(int i, int j, int k) const { return (static_cast<long>(k) * ny + j) * nx + i; }
  long nid(int i, int j, int k) const { return (static_cast<long>(k) * (ny + 1) + j) * (nx + 1) + i; }
  bool is_active(int i, int j, int k) const {
    if (i < 0 || j < 0 || k < 0 || i >= nx || j >= ny || k >= nz) return false;
    return active[cid(i, j, k)] != 0;
  }
};

struct BoundaryQuad {
  int i, j, k, axis, side;  // owning cell and which of its faces
};

// Tag lookup for a boundary face: callback decides the physical tag.
using TagFn = int (*)(const void* ctx, const Lattice& L, const BoundaryQuad& q);

ElementStream build(const Lattice& L, double jitter, uint64_t seed, const void* tag_ctx, TagFn tag_fn,
                    int volume_tag) {
  const long nn = static_cast<long>(L.nx + 1) * (L.ny + 1) * (L.nz + 1);
  // boundary quads, per side in order x-,x+,y-,y+,z-,z+ ; lexicographic within a side
  std::vector<BoundaryQuad> bq;
  for (int axis = 0; axis < 3; ++axis)
    for (int side = 0; side < 2; ++side)
      for (int k = 0; k < L.nz; ++k)
        for (int j = 0; j < L.ny; ++j)
          for (int i = 0; i < L.nx; ++i) {
            if (!L.is_active(i, j, k)) continue;
            int d[3] = {0, 0, 0};
            d[axis] = side ? 1 : -1;
            if (!L.is_active(i + d[0], j + d[1], k + d[2])) bq.push_back({i, j, k, axis, side});
          }
  // node usage and fixed axes (a node on a boundary face may not move along its normal)
  std::vector<uint8_t> used(nn, 0), fixed(nn, 0);
  auto corner_node = [&](int i, int j, int k, int c) {
    return L.nid(i + (c & 1), j + ((c >> 1) & 1), k + ((c >> 2) & 1));
  };
  for (int k = 0; k < L.nz; ++k)
    for (int j = 0; j < L.ny; ++j)
      for (int i = 0; i < L.nx; ++i)
        if (L.is_active(i, j, k))
          for (int c = 0; c < 8; ++c) used[corner_node(i, j, k, c)] = 1;
  for (const auto& q : bq)
    for (int c = 0; c < 8; ++c) {
      int bit = (c >> q.axis) & 1;
      if (bit == q.side) fixed[corner_node(q.i, q.j, q.k, c)] |= static_cast<uint8_t>(1u << q.axis);
    }
  // compact numbering of used nodes, lexicographic
  std::vector<int> newid(nn, -1);
  ElementStream es;
  Rng rng(seed);
  int nv = 0;
  for (int k = 0; k <= L.nz; ++k)
    for (int j = 0; j <= L.ny; ++j)
      for (int i = 0; i <= L.nx; ++i) {
        long n = L.nid(i, j, k);
        if (!used[n]) continue;
        newid[n] = nv++;
        double p[3] = {L.xs[i], L.ys[j], L.zs[k]};
        const int idx[3] = {i, j, k};
        const std::vector<double>* ax[3] = {&L.xs, &L.ys, &L.zs};
        // draw three numbers for every node so the field is independent of masks
        double r[3] = {rng.sym(), rng.sym(), rng.sym()};
        if (jitter > 0.0)
          for (int a = 0; a < 3; ++a) {
            if (fixed[n] & (1u << a)) continue;
            const auto& v = *ax[a];
            int m = idx[a];
            double hl = m > 0 ? v[m] - v[m - 1] : v[m + 1] - v[m];
            double hr = m + 1 < static_cast<int>(v.size()) ? v[m + 1] - v[m] : hl;
            p[a] += jitter * std::min(hl, hr) * r[a];
          }
        es.xyz.push_back(p[0]);
        es.xyz.push_back(p[1]);
        es.xyz.push_back(p[2]);
      }
  // boundary elements
  static const int kOther[3][2] = {{1, 2}, {2, 0}, {0, 1}};
  for (const auto& q : bq) {
    int p = kOther[q.axis][0], r = kOther[q.axis][1];
    // face corners (cyclic): (p,r) = (0,0),(1,0),(1,1),(0,1)
    int cc[4];
    const int pr[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
    for (int t = 0; t < 4; ++t) cc[t] = (q.side << q.axis) | (pr[t][0] << p) | (pr[t][1] << r);
    int v[4];
    for (int t = 0; t < 4; ++t) v[t] = newid[corner_node(q.i, q.j, q.k, cc[t])];
    int tagv = tag_fn(tag_ctx, L, q);
    if (!L.split[L.cid(q.i, q.j, q.k)]) {
      es.add(kQuad, tagv, v);
    } else {
      // Kuhn diagonal: through corner (000) on a low face (t=0), through (111) on a high face (t=2)
      if (q.side == 0) {
        int t1[3] = {v[0], v[1], v[2]}, t2[3] = {v[0], v[2], v[3]};
        es.add(kTri, tagv, t1);
        es.add(kTri, tagv, t2);
      } else {
        // high face: corner with p=r=1 is (111) -> diagonal v[2]-v[0]
        int t1[3] = {v[2], v[3], v[0]}, t2[3] = {v[2], v[0], v[1]};
        es.add(kTri, tagv, t1);
        es.add(kTri, tagv, t2);
      }
    }
  }
  // volume elements
  static const int kPerm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  for (int k = 0; k < L.nz; ++k)
    for (int j = 0; j < L.ny; ++j)
      for (int i = 0; i < L.nx; ++i) {
        if (!L.is_active(i, j, k)) continue;
        if (!L.split[L.cid(i, j, k)]) {
          int v[8];
          for (int c = 0; c < 8; ++c) v[c] = newid[corner_node(i, j, k, kHexCorner[c])];
          es.add(kHex, volume_tag, v);
        } else {
          for (const auto& s : kPerm) {
            int c1 = 1 << s[0], c2 = c1 | (1 << s[1]);
            int v[4] = {newid[corner_node(i, j, k, 0)], newid[corner_node(i, j, k, c1)],
                        newid[corner_node(i, j, k, c2)], newid[corner_node(i, j, k, 7)]};
            es.add(kTet, volume_tag, v);
          }
        }
      }
  return es;
}

std::vector<double> axis_nodes(double x0, double len, int n) {
  std::vector<double> v(n + 1);
  for (int i = 0; i <= n; ++i) v[i] = x0 + len * static_cast<double>(i) / n;
  return v;
}

int box_tag(const void* ctx, const Lattice&, const BoundaryQuad& q) {
  const BoxSpec* s = static_cast<const BoxSpec*>(ctx);
  return s->side_tag[2 * q.axis + q.side];
}

}  // namespace

ElementStream make_box(const BoxSpec& spec) {
  if (spec.nx < 1 || spec.ny < 1 || spec.nz < 1) throw std::invalid_argument("box: empty lattice");
  if (spec.jitter < 0.0 || spec.jitter >= 0.25) throw std::invalid_argument("box: jitter must be in [0,0.25)");
  Lattice L;
  L.nx = spec.nx;
  L.ny = spec.ny;
  L.nz = spec.nz;
  L.xs = axis_nodes(spec.x0, spec.lx, spec.nx);
  L.ys = axis_nodes(spec.y0, spec.ly, spec.ny);
  L.zs = axis_nodes(spec.z0, spec.lz, spec.nz);
  long nc = static_cast<long>(L.nx) * L.ny * L.nz;
  L.active.assign(nc, 1);
  L.split.assign(nc, spec.all_tets ? 1 : 0);
  if (!spec.all_tets && spec.split_frac > 0.0) {
    Rng r(spec.seed ^ 0x5bd1e995ull);
    for (long c = 0; c < nc; ++c) L.split[c] = r.uniform() < spec.split_frac ? 1 : 0;
  }
  return build(L, spec.jitter, spec.seed, &spec, box_tag, spec.volume_tag);
}

namespace {
struct StepCtx {
  const StepSpec* s;
  int n_up;  // cells upstream of the step edge along x
  int n_h;   // cells per step height
};
int step_tag(const void* ctx, const Lattice& L, const BoundaryQuad& q) {
  const StepCtx* c = static_cast<const StepCtx*>(ctx);
  if (q.axis == 0 && q.side == 0 && q.i == 0) return c->s->tag_inlet;
  if (q.axis == 0 && q.side == 1 && q.i == L.nx - 1) return c->s->tag_outlet;
  if (q.axis == 2) return c->s->tag_sym;
  return c->s->tag_wall;
}
}  // namespace

ElementStream make_step(const StepSpec& spec) {
  const int n = spec.n_per_h;
  if (n < 1) throw std::invalid_argument("step: n_per_h < 1");
  StepCtx ctx{&spec, static_cast<int>(std::lround(spec.upstream / spec.h * n)), n};
  Lattice L;
  L.nx = ctx.n_up + static_cast<int>(std::lround(spec.downstream / spec.h * n));
  L.ny = 2 * n;
  L.nz = static_cast<int>(std::lround(spec.span / spec.h * n));
  L.xs = axis_nodes(-spec.upstream, spec.upstream + spec.downstream, L.nx);
  L.ys = axis_nodes(0.0, 2.0 * spec.h, L.ny);
  L.zs = axis_nodes(0.0, spec.span, L.nz);
  long nc = static_cast<long>(L.nx) * L.ny * L.nz;
  L.active.assign(nc, 1);
  L.split.assign(nc, spec.tets ? 1 : 0);
  for (int k = 0; k < L.nz; ++k)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < ctx.n_up; ++i) L.active[L.cid(i, j, k)] = 0;
  return build(L, spec.jitter, spec.seed, &ctx, step_tag, spec.volume_tag);
}

}  // namespace s355
