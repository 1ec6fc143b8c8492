// Device-side views of the mesh and the flow fields.  Plain structs of raw
// device pointers, passed to kernels by value.
#pragma once
#include <hip/hip_runtime.h>

#include "common.hpp"

namespace s355 {

// Boundary type codes (BdRegion::type1)
enum : int { kWall = 1, kInlet = 2, kOutlet = 3, kSym = 4 };

struct DMesh {
  int Nc = 0, Nf = 0, Nb = 0, nnz = 0;
  // faces; faces 0..Nb-1 are the boundary faces, face b <-> boundary record b
  const int* f_c1 = nullptr;
  const int* f_c2 = nullptr;
  const double* f_lambda = nullptr;
  const double* f_rlencos = nullptr;
  const double* f_area = nullptr;
  const double* f_n = nullptr;     // 3 per face
  const double* f_x = nullptr;     // 3 per face
  const double* f_xpac = nullptr;  // 3 per face
  const double* f_xnac = nullptr;  // 3 per face
  // cells
  const double* c_vol = nullptr;
  const double* c_x = nullptr;  // 3 per cell
  const int* cf_ptr = nullptr;
  const int* cf_pat = nullptr;
  // half-face streams: every cell's entries, cells in processing order
  // (c_order), one record {cell, face, neighbour code, CSR slot (-1 on a
  // boundary face)} per entry, the neighbour code being c2 (>= 0) where the
  // cell is the face's c1, ~c1 (< 0) where it is c2, and 0 on boundary faces;
  // hf in pattern order (the assembly loops), hfs in ascending face order
  // (the gradient).  Block b of a half-face kernel takes entries
  // [blk[b], blk[b + 1]): at most kBlock of them, whole cells only, at most
  // kWave cells.
  const int4* hf = nullptr;
  const int* hf_blk = nullptr;
  int hf_nblk = 0;
  const int4* hfs = nullptr;
  const int* hfs_blk = nullptr;
  int hfs_nblk = 0;
  // CSR
  const int* rowptr = nullptr;
  const int* col = nullptr;
  const int* diag = nullptr;
  // boundary records
  const int* b_type = nullptr;
  const double* b_dist = nullptr;
  // independent cases stored one after the other (one for a single case):
  // kBlock-row tiles of the cells and of the boundary faces, for the per-case
  // reductions (residuals, outlet mass flow, pressure reference)
  Tiles ctile, btile;
  // processing orders of the cell and face loops: thread t of a cell kernel
  // takes cell c_order[t], of a face kernel face f_order[t] -- the cells in
  // the order of the input elements (spatially coherent, where the partition
  // sort's numbering is not), the interior faces by their first cell's
  // place in it.  Results do not depend on them (one writer per output).
  const int* c_order = nullptr;
  const int* f_order = nullptr;
  // cells touching a wall, with their wall faces in ascending order
  int n_wcell = 0;
  const int* wc_cell = nullptr;
  const int* wc_ptr = nullptr;
  const int* wc_face = nullptr;
  // vertices, for the gradient limiters (uploaded when a limiter is on):
  // each cell's 8 vertices (degenerate hex), coordinates, and the cells of
  // each vertex (ascending, one entry per cell)
  int Nv = 0;
  const int* c_vert = nullptr;
  const double* vert = nullptr;
  const int* v2c_ptr = nullptr;
  const int* v2c = nullptr;
};

struct DFields {
  double *R, *U, *V, *W, *P, *T, *TE, *ED, *dp, *mu, *mut, *apr;
  double *dUdX, *dVdX, *dWdX, *dPdX, *dPhidX;
  double *BRo, *BU, *BV, *BW, *BPre, *BTE, *BED, *BTem, *shear;
  double* RUFace;
  double *Au, *Ap, *As, *bu, *bv, *bw, *bp, *bs;
  // old time levels (unsteady runs; null otherwise): *np the previous step,
  // *np2 the one before (SaveTransientOldData, navier.cpp:700-742)
  double *Unp, *Vnp, *Wnp, *Tnp, *TEp, *EDp;
  double *Unp2, *Vnp2, *Wnp2, *Tnp2, *TEp2, *EDp2;
};

}  // namespace s355
