// Mesh geometry on the device: CellFaceInfo's cell and face metrics
// (geometry.cpp:391-752) from the vertices and the topology the host ingest
// built.  Every expression is evaluated in the host ingest's order (built
// with -ffp-contract=off; sqrt and division are correctly rounded on both
// sides), so the device metrics equal the host's bit for bit.
#include "common.hpp"
#include "geometry.hpp"

namespace s355 {
namespace {

__device__ __forceinline__ void sub3(double* r, const double* a, const double* b) {
  r[0] = a[0] - b[0];
  r[1] = a[1] - b[1];
  r[2] = a[2] - b[2];
}

__device__ __forceinline__ double dot3(const double* a, const double* b) {
  double s = 0.;
  s += a[0] * b[0];
  s += a[1] * b[1];
  s += a[2] * b[2];
  return s;
}

__device__ __forceinline__ double len3(const double* a) { return sqrt(dot3(a, a)); }

__device__ __forceinline__ void cross3(double* a, const double* b, const double* c) {
  a[0] = b[1] * c[2] - b[2] * c[1];
  a[1] = -b[0] * c[2] + b[2] * c[0];
  a[2] = b[0] * c[1] - b[1] * c[0];
}

// area of the triangle x1 x2 x3
__device__ double tri_area(const double* x1, const double* x2, const double* x3) {
  double v1[3], v2[3];
  sub3(v1, x2, x1);
  sub3(v2, x3, x1);
  const double a1 = v1[1] * v2[2] - v1[2] * v2[1];
  const double a2 = -v1[0] * v2[2] + v1[2] * v2[0];
  const double a3 = v1[0] * v2[1] - v1[1] * v2[0];
  return 0.5 * sqrt(a1 * a1 + a2 * a2 + a3 * a3);
}

// volume of the tetrahedron x1 x2 x3 x4
__device__ double tet_vol(const double* x1, const double* x2, const double* x3, const double* x4) {
  double v1[3], d1[3], d2[3], d3[3];
  sub3(d1, x2, x1);
  sub3(d2, x3, x1);
  sub3(d3, x4, x1);
  cross3(v1, d1, d2);
  return 1. / 6. * fabs(dot3(v1, d3));
}

__device__ __forceinline__ void load3(double* r, const double* x, int v) {
  r[0] = x[3 * static_cast<size_t>(v)];
  r[1] = x[3 * static_cast<size_t>(v) + 1];
  r[2] = x[3 * static_cast<size_t>(v) + 2];
}

// Cell volume and centre.  Hexahedra (degenerate ones included): the vertex
// mean, then the volume as 6 tetrahedra and the volume-weighted centre of
// their centroids (geometry.cpp:412-452).  Cells whose hex quads are split
// into triangles of split neighbours: tetrahedra on the face triangles about
// the vertex mean.
__global__ void __launch_bounds__(kBlock) k_cell_metrics(GeomArgs g) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.Nc) return;
  const int* h = &g.c_vert[8 * static_cast<size_t>(i)];
  double x[8][3];
#pragma unroll
  for (int k = 0; k < 8; ++k) load3(x[k], g.vert, h[k]);
  double xc[3];
#pragma unroll
  for (int j = 0; j < 3; ++j)
    xc[j] = 1. / 8. * (x[0][j] + x[1][j] + x[2][j] + x[3][j] + x[4][j] + x[5][j] + x[6][j] + x[7][j]);
  double* oc = &g.c_x[3 * static_cast<size_t>(i)];
  if (g.c_poly[i]) {
    const double xp[3] = {xc[0], xc[1], xc[2]};
    double vs = 0., acc[3] = {0., 0., 0.};
    for (int e = g.cf_ptr[i]; e < g.cf_ptr[i + 1]; ++e) {
      const int* q = &g.f_v[4 * static_cast<size_t>(g.cf_pat[e])];
      const int tri[2][3] = {{q[0], q[1], q[3]}, {q[1], q[2], q[3]}};
      for (int t = 0; t < 2; ++t) {
        double p0[3], p1[3], p2[3];
        load3(p0, g.vert, tri[t][0]);
        load3(p1, g.vert, tri[t][1]);
        load3(p2, g.vert, tri[t][2]);
        const double v = tet_vol(xp, p0, p1, p2);
        vs += v;
        for (int k = 0; k < 3; ++k) acc[k] += v * 0.25 * (xp[k] + p0[k] + p1[k] + p2[k]);
      }
    }
    g.c_vol[i] = vs;
    for (int k = 0; k < 3; ++k) oc[k] = acc[k] / vs;
    return;
  }
  const double *x1 = x[0], *x2 = x[1], *x3 = x[2], *x4 = x[3], *x5 = x[4], *x6 = x[5], *x7 = x[6], *x8 = x[7];
  const double V1 = tet_vol(x1, x2, x4, x6), V2 = tet_vol(x1, x5, x4, x6), V3 = tet_vol(x4, x5, x8, x6);
  const double V4 = tet_vol(x2, x3, x4, x7), V5 = tet_vol(x2, x6, x4, x7), V6 = tet_vol(x8, x4, x6, x7);
  g.c_vol[i] = V1 + V2 + V3 + V4 + V5 + V6;
  const double vs = V1 + V2 + V3 + V4 + V5 + V6;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const double c0 = 0.25 * (x1[j] + x2[j] + x4[j] + x6[j]);
    const double c1 = 0.25 * (x1[j] + x5[j] + x4[j] + x6[j]);
    const double c2 = 0.25 * (x4[j] + x5[j] + x8[j] + x6[j]);
    const double c3 = 0.25 * (x2[j] + x3[j] + x4[j] + x7[j]);
    const double c4 = 0.25 * (x2[j] + x6[j] + x4[j] + x7[j]);
    const double c5 = 0.25 * (x8[j] + x4[j] + x6[j] + x7[j]);
    oc[j] = 1. / vs * (V1 * c0 + V2 * c1 + V3 * c2 + V4 * c3 + V5 * c4 + V6 * c5);
  }
}

// Face centre (area-weighted centroids of the triangles 124 and 234), area
// vector (half the cross product of the diagonals, pointing from c1 out),
// area, interpolation weight, rlencos and the projections Xpac / Xnac of
// the cell centres onto the face normal line (geometry.cpp:560-752); the
// wall distance of each boundary face; warped[f] = 1 where the two triangle
// areas differ from the area vector's length by more than 1e-4.
__global__ void __launch_bounds__(kBlock) k_face_metrics(GeomArgs g) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= g.Nf) return;
  const int* q = &g.f_v[4 * static_cast<size_t>(f)];
  double p1[3], p2[3], p3[3], p4[3];
  load3(p1, g.vert, q[0]);
  load3(p2, g.vert, q[1]);
  load3(p3, g.vert, q[2]);
  load3(p4, g.vert, q[3]);
  const double a1 = tri_area(p1, p2, p4), a2 = tri_area(p2, p3, p4);
  double xc1[3], xc2[3], fx[3], fn[3], d1[3], d2[3], dx[3];
  for (int j = 0; j < 3; ++j) {
    xc1[j] = 1. / 3 * (p1[j] + p2[j] + p4[j]);
    xc2[j] = 1. / 3 * (p2[j] + p3[j] + p4[j]);
  }
  for (int j = 0; j < 3; ++j) fx[j] = (a1 * xc1[j] + a2 * xc2[j]) / (a1 + a2);
  sub3(d1, p1, p3);
  sub3(d2, p2, p4);
  cross3(fn, d1, d2);
  for (int j = 0; j < 3; ++j) fn[j] = 0.5 * fn[j];
  const int c1 = g.f_c1[f];
  double x1[3];
  load3(x1, g.c_x, c1);
  sub3(dx, fx, x1);
  if (dot3(fn, dx) < 0.)
    for (int j = 0; j < 3; ++j) fn[j] = -fn[j];
  const double area = len3(fn);
  g.warped[f] = fabs((a1 + a2 - area) / area) > 1.e-4 ? 1 : 0;
  double* ox = &g.f_x[3 * static_cast<size_t>(f)];
  double* on = &g.f_n[3 * static_cast<size_t>(f)];
  double* opa = &g.f_xpac[3 * static_cast<size_t>(f)];
  double* ona = &g.f_xnac[3 * static_cast<size_t>(f)];
  for (int j = 0; j < 3; ++j) {
    ox[j] = fx[j];
    on[j] = fn[j];
  }
  g.f_area[f] = area;
  const int c2 = g.f_c2[f];
  if (c2 >= 0) {
    double x2[3];
    load3(x2, g.c_x, c2);
    sub3(d1, fx, x1);
    sub3(d2, fx, x2);
    const double r1 = len3(d1), r2 = len3(d2);
    g.f_lambda[f] = r2 / (r1 + r2);
    sub3(dx, x2, x1);
    g.f_rlencos[f] = area / fabs(dot3(dx, fn) / area);
    sub3(dx, fx, x1);
    double ss = dot3(dx, fn) / area;
    for (int j = 0; j < 3; ++j) dx[j] = ss * fn[j] / area;
    sub3(opa, fx, dx);
    sub3(dx, fx, x2);
    ss = dot3(dx, fn) / area;
    for (int j = 0; j < 3; ++j) dx[j] = ss * fn[j] / area;
    sub3(ona, fx, dx);
  } else {
    g.f_lambda[f] = 1.;
    sub3(dx, fx, x1);
    g.f_rlencos[f] = area / (dot3(dx, fn) / area);
    for (int j = 0; j < 3; ++j) opa[j] = ona[j] = 0.0;
  }
  if (f < g.Nb) {
    // the wall distance of boundary record f (faces 0..Nb-1 are the boundary)
    sub3(dx, fx, x1);
    g.b_dist[f] = dot3(dx, fn) / area;
  }
}

}  // namespace

void launch_geometry(const GeomArgs& g, hipStream_t s) {
  hipLaunchKernelGGL(k_cell_metrics, dim3(grid_for(g.Nc)), dim3(kBlock), 0, s, g);
  S355_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_face_metrics, dim3(grid_for(g.Nf)), dim3(kBlock), 0, s, g);
  S355_LAUNCH_CHECK();
}

}  // namespace s355
