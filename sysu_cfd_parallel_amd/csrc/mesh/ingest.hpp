// Mesh ingest: element stream -> cells, boundary records, faces, metrics and
// the CSR pattern, in the reference's numbering.  Arrays are structure-of-
// arrays so they upload to HBM unchanged.
#pragma once
#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "../../../include/simple355.h"

namespace s355 {

constexpr int kMaxFace = S355_MAX_CELL_FACES;

// Allocator whose resize() leaves elements uninitialised, so that multi-GB
// mesh arrays are first touched by the parallel loops that fill them.
template <class T>
struct DefaultInit : std::allocator<T> {
  template <class U>
  struct rebind {
    using other = DefaultInit<U>;
  };
  DefaultInit() = default;
  template <class U>
  DefaultInit(const DefaultInit<U>&) noexcept {}
  template <class U>
  void construct(U* p) noexcept {
    ::new (static_cast<void*>(p)) U;
  }
  template <class U, class... A>
  void construct(U* p, A&&... a) {
    ::new (static_cast<void*>(p)) U(std::forward<A>(a)...);
  }
};
template <class T>
using hvec = std::vector<T, DefaultInit<T>>;
constexpr int kVoidBoundary = -10;

struct Mesh {
  int n_vert = 0, n_cell = 0, n_face = 0, n_bnd = 0, n_elem = 0, n_warped = 0;
  hvec<int> perm;      // element order after the partition sort
  hvec<double> vert;   // 3 per vertex

  // faces: boundary faces are 0..n_bnd-1 with face i <-> boundary record i
  hvec<int> f_c1, f_c2, f_bnd, f_v;  // f_v: 4 per face
  hvec<double> f_x, f_n, f_xpac, f_xnac;  // 3 per face
  hvec<double> f_area, f_lambda, f_rlencos;

  // cells: faces in the reference's pattern order (CellData::face)
  hvec<int> c_rid, c_nface, c_face, c_cell;  // c_face/c_cell: kMaxFace per cell, -1 padded
  hvec<int> c_vert;  // 8 per cell: the element's vertices as a degenerate hex (CellData::vertices)
  hvec<int> c_elem;  // the element each cell comes from (its index in the input stream)
  hvec<uint8_t> c_poly;
  hvec<double> c_vol, c_x;  // c_x: 3 per cell

  // boundary records
  hvec<int> b_rid;
  hvec<double> b_dist;

  // CSR pattern (rows = cells, columns ascending like PETSc AIJ)
  hvec<int> rowptr, col, diag;
  hvec<int> slot;  // kMaxFace per cell: CSR index of the neighbour, -1 for boundary

  // compact cell->face lists for device gathers
  hvec<int> cf_ptr;     // n_cell + 1
  hvec<int> cf_pat;     // faces in pattern order
  hvec<int> cf_slot;    // CSR slot per pattern entry (-1 boundary)
  hvec<int> cf_sorted;  // faces in ascending order

  std::vector<s355_region> regions;

  int nnz() const { return static_cast<int>(col.size()); }
  int type1_of_rid(int rid) const;
};

// Builds the mesh; throws std::runtime_error with a message on bad input.
// metrics = false leaves the geometry (cell volumes and centres, face
// metrics, wall distances, the warped-face count) empty: the device context
// computes it on the GPU from the same topology and vertices.
Mesh ingest_mesh(const s355_mesh_desc& desc, bool metrics = true);

// Offsets of the cases of a batch mesh (ncase + 1 entries each): case k owns
// cells [cell[k], cell[k+1]), boundary records [bnd[k], bnd[k+1]), interior
// faces [n_bnd + iface[k], n_bnd + iface[k+1]), CSR entries, vertices,
// elements and cell-face list entries likewise; its region ids are shifted
// by k * rid_stride.
struct BatchLayout {
  std::vector<int> cell, bnd, iface, nnz, vert, elem, cf;
  int rid_stride = 1;
};

// Concatenates independent cases into one mesh: cells, rows and vertices
// case after case; the boundary faces of all cases first, then their
// interior faces, each case keeping its own face order.
Mesh concat_meshes(const std::vector<Mesh>& parts, BatchLayout& lay);

}  // namespace s355
