// Batch layout: independent cases concatenated into one mesh so that every
// kernel covers all of them in one launch.
//
// Cells, CSR rows and vertices of case k follow those of case k-1.  Faces are
// renumbered so that the boundary faces of all cases come first (case by
// case, each in its own order), then the interior faces of all cases (case by
// case, each in its own order): "face f is boundary record f" holds for the
// whole batch, and inside one case the order of its faces, hence of every
// per-cell face list, is unchanged.  Region ids of case k are shifted by
// k * stride so that the cases' region tables stay apart.
#include <algorithm>
#include <stdexcept>

#include "ingest.hpp"
#include "parallel.hpp"

namespace s355 {

namespace {
template <class T>
void append(hvec<T>& dst, const hvec<T>& src) {
  dst.insert(dst.end(), src.begin(), src.end());
}
}  // namespace

Mesh concat_meshes(const std::vector<Mesh>& parts, BatchLayout& lay) {
  const int nc = static_cast<int>(parts.size());
  if (nc < 1) throw std::invalid_argument("batch: no cases");
  lay = BatchLayout{};
  lay.cell.assign(nc + 1, 0);
  lay.bnd.assign(nc + 1, 0);
  lay.iface.assign(nc + 1, 0);
  lay.nnz.assign(nc + 1, 0);
  lay.vert.assign(nc + 1, 0);
  lay.elem.assign(nc + 1, 0);
  lay.cf.assign(nc + 1, 0);
  int max_rid = 0;
  for (int k = 0; k < nc; ++k) {
    const Mesh& p = parts[k];
    if (p.n_cell < 1) throw std::invalid_argument("batch: case " + std::to_string(k) + " has no cells");
    for (const auto& r : p.regions) max_rid = std::max(max_rid, r.rid);
    const long cells = static_cast<long>(lay.cell[k]) + p.n_cell;
    const long bnds = static_cast<long>(lay.bnd[k]) + p.n_bnd;
    const long ifs = static_cast<long>(lay.iface[k]) + (p.n_face - p.n_bnd);
    const long nnz = static_cast<long>(lay.nnz[k]) + p.nnz();
    const long verts = static_cast<long>(lay.vert[k]) + p.n_vert;
    const long elems = static_cast<long>(lay.elem[k]) + p.n_elem;
    const long cfs = static_cast<long>(lay.cf[k]) + p.cf_ptr[p.n_cell];
    const long lim = 0x7fffffffL;
    if (cells > lim || bnds + ifs > lim || nnz > lim || verts > lim || elems > lim || cfs > lim)
      throw std::invalid_argument("batch: total size above 32-bit indices");
    lay.cell[k + 1] = static_cast<int>(cells);
    lay.bnd[k + 1] = static_cast<int>(bnds);
    lay.iface[k + 1] = static_cast<int>(ifs);
    lay.nnz[k + 1] = static_cast<int>(nnz);
    lay.vert[k + 1] = static_cast<int>(verts);
    lay.elem[k + 1] = static_cast<int>(elems);
    lay.cf[k + 1] = static_cast<int>(cfs);
  }
  lay.rid_stride = max_rid + 1;

  Mesh m;
  m.n_cell = lay.cell[nc];
  m.n_bnd = lay.bnd[nc];
  m.n_face = m.n_bnd + lay.iface[nc];
  m.n_vert = lay.vert[nc];
  m.n_elem = lay.elem[nc];
  const int Nb = m.n_bnd;
  auto gface = [&](int k, int f) {
    const Mesh& p = parts[k];
    return f < p.n_bnd ? lay.bnd[k] + f : Nb + lay.iface[k] + (f - p.n_bnd);
  };
  const size_t Nf = m.n_face, Nc = m.n_cell, K = kMaxFace;
  // the parts' geometry, when they carry it (ingest_mesh with metrics)
  const bool met = !parts[0].c_vol.empty();
  for (const Mesh& p : parts)
    if (p.c_vol.empty() == met) throw std::invalid_argument("batch: parts with and without geometry");
  m.f_c1.resize(Nf);
  m.f_c2.resize(Nf);
  m.f_bnd.resize(Nf);
  m.f_v.resize(4 * Nf);
  if (met) {
    m.f_x.resize(3 * Nf);
    m.f_n.resize(3 * Nf);
    m.f_xpac.resize(3 * Nf);
    m.f_xnac.resize(3 * Nf);
    m.f_area.resize(Nf);
    m.f_lambda.resize(Nf);
    m.f_rlencos.resize(Nf);
    m.b_dist.resize(Nb);
  }
  m.c_face.resize(Nc * K);
  m.c_cell.resize(Nc * K);
  m.slot.resize(Nc * K);
  m.b_rid.resize(Nb);
  m.rowptr.resize(Nc + 1);
  m.rowptr[0] = 0;
  m.col.resize(lay.nnz[nc]);
  m.diag.resize(Nc);
  m.cf_ptr.resize(Nc + 1);
  m.cf_ptr[0] = 0;
  m.cf_pat.resize(lay.cf[nc]);
  m.cf_slot.resize(lay.cf[nc]);
  m.cf_sorted.resize(lay.cf[nc]);
  parallel_for(nc, [&](long kl) {
    const int k = static_cast<int>(kl);
    const Mesh& p = parts[k];
    const int C0 = lay.cell[k], B0 = lay.bnd[k], Z0 = lay.nnz[k], V0 = lay.vert[k], Q0 = lay.cf[k];
    const int R0 = k * lay.rid_stride;
    auto cell = [&](int c) { return c >= 0 ? c + C0 : c; };
    for (int f = 0; f < p.n_face; ++f) {
      const size_t g = gface(k, f);
      m.f_c1[g] = cell(p.f_c1[f]);
      m.f_c2[g] = cell(p.f_c2[f]);
      m.f_bnd[g] = p.f_bnd[f] >= 0 ? p.f_bnd[f] + B0 : p.f_bnd[f];
      for (int j = 0; j < 4; ++j) {
        const int v = p.f_v[4 * static_cast<size_t>(f) + j];
        m.f_v[4 * g + j] = v >= 0 ? v + V0 : v;
      }
      if (!met) continue;
      for (int j = 0; j < 3; ++j) {
        m.f_x[3 * g + j] = p.f_x[3 * static_cast<size_t>(f) + j];
        m.f_n[3 * g + j] = p.f_n[3 * static_cast<size_t>(f) + j];
        m.f_xpac[3 * g + j] = p.f_xpac[3 * static_cast<size_t>(f) + j];
        m.f_xnac[3 * g + j] = p.f_xnac[3 * static_cast<size_t>(f) + j];
      }
      m.f_area[g] = p.f_area[f];
      m.f_lambda[g] = p.f_lambda[f];
      m.f_rlencos[g] = p.f_rlencos[f];
    }
    for (int b = 0; b < p.n_bnd; ++b) {
      m.b_rid[B0 + b] = p.b_rid[b] + R0;
      if (met) m.b_dist[B0 + b] = p.b_dist[b];
    }
    for (int i = 0; i < p.n_cell; ++i) {
      const size_t gi = static_cast<size_t>(C0 + i);
      for (size_t j = 0; j < K; ++j) {
        const size_t l = static_cast<size_t>(i) * K + j;
        const int f = p.c_face[l];
        m.c_face[gi * K + j] = f >= 0 ? gface(k, f) : f;
        m.c_cell[gi * K + j] = cell(p.c_cell[l]);
        m.slot[gi * K + j] = p.slot[l] >= 0 ? p.slot[l] + Z0 : p.slot[l];
      }
      m.rowptr[gi + 1] = p.rowptr[i + 1] + Z0;
      m.diag[gi] = p.diag[i] + Z0;
      m.cf_ptr[gi + 1] = p.cf_ptr[i + 1] + Q0;
    }
    for (int z = 0; z < p.nnz(); ++z) m.col[Z0 + z] = p.col[z] + C0;
    for (int q = 0; q < p.cf_ptr[p.n_cell]; ++q) {
      m.cf_pat[Q0 + q] = gface(k, p.cf_pat[q]);
      m.cf_sorted[Q0 + q] = gface(k, p.cf_sorted[q]);
      m.cf_slot[Q0 + q] = p.cf_slot[q] >= 0 ? p.cf_slot[q] + Z0 : p.cf_slot[q];
    }
  });
  for (int k = 0; k < nc; ++k) {
    const Mesh& p = parts[k];
    const int R0 = k * lay.rid_stride;
    for (int e : p.perm) m.perm.push_back(e + lay.elem[k]);
    append(m.vert, p.vert);
    for (int r : p.c_rid) m.c_rid.push_back(r + R0);
    for (int v : p.c_vert) m.c_vert.push_back(v + lay.vert[k]);
    for (int e : p.c_elem) m.c_elem.push_back(e + lay.elem[k]);
    append(m.c_nface, p.c_nface);
    append(m.c_poly, p.c_poly);
    append(m.c_vol, p.c_vol);
    append(m.c_x, p.c_x);
    for (s355_region r : p.regions) {
      r.rid += R0;
      m.regions.push_back(r);
    }
    m.n_warped += p.n_warped;
  }
  return m;
}

}  // namespace s355
