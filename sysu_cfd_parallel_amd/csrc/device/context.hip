// Device context: owns the HBM copies of one case and runs the device-level
// operations of the SIMPLE hot path on one HIP stream.
#include "context.hpp"
#include "geometry.hpp"
#include "scratch.hpp"
#include "../mesh/reorder.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "common.hpp"

namespace s355 {

namespace {
constexpr double kSmallInit = 1.e-16;
constexpr double kCmuInit = 0.09;
}  // namespace

template <class T>
T* Context::alloc(size_t n) {
  void* p = nullptr;
  size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
  S355_HIP(hipMalloc(&p, bytes));
  allocs_.push_back(p);
  return static_cast<T*>(p);
}

namespace {
template <class Vec>
auto upload(std::vector<void*>& allocs, const Vec& v, hipStream_t s) {
  using T = typename Vec::value_type;
  void* p = nullptr;
  S355_HIP(hipMalloc(&p, std::max<size_t>(v.size(), 1) * sizeof(T)));
  allocs.push_back(p);
  if (!v.empty()) S355_HIP(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
  return static_cast<const T*>(p);
}
}  // namespace

Context::Context(const s355_mesh_desc* descs, int n_case, const s355_params& params) : par(params), ncase(n_case) {
  if (par.density_model != 0) throw std::invalid_argument("only DensityModel 0 (constant density) is supported");
  if (par.precision != S355_FP64 && par.precision != S355_MIXED) throw std::invalid_argument("bad precision mode");
  if (par.inner_order < 0 || par.inner_order > 2) throw std::invalid_argument("inner_order: 0, 1 (RCM) or 2 (input)");
  if (par.grad_scheme < 0 || par.grad_scheme > 1) throw std::invalid_argument("grad_scheme: 0 Green-Gauss, 1 least squares");
  if (par.assembly_fp32 != 0 && par.assembly_fp32 != 1) throw std::invalid_argument("assembly_fp32: 0 or 1");
  if (!descs || ncase < 1) throw std::invalid_argument("no mesh");
  if (ncase > 1)
    for (int k = 0; k < 3; ++k)
      if (par.ksp[k].type == S355_KSP_CG)
        throw std::invalid_argument("a batch of cases solves every system with GMRES (KSP type CG requested)");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) throw HipError("no HIP device available");
  if (ncase == 1) {
    mesh = ingest_mesh(descs[0], false);  // the geometry is computed on the device
    lay.cell = {0, mesh.n_cell};
    lay.bnd = {0, mesh.n_bnd};
    lay.iface = {0, mesh.n_face - mesh.n_bnd};
    lay.nnz = {0, mesh.nnz()};
    lay.vert = {0, mesh.n_vert};
    lay.elem = {0, mesh.n_elem};
    lay.cf = {0, mesh.cf_ptr[mesh.n_cell]};
    lay.rid_stride = 0;
  } else {
    std::vector<Mesh> parts;
    parts.reserve(ncase);
    for (int k = 0; k < ncase; ++k) parts.push_back(ingest_mesh(descs[k], false));
    mesh = concat_meshes(parts, lay);
  }
  int min_cells = mesh.n_cell;
  for (int k = 0; k < ncase; ++k) min_cells = std::min(min_cells, lay.cell[k + 1] - lay.cell[k]);
  if (par.cell_pressure_ref < 0 || par.cell_pressure_ref >= min_cells)
    throw std::invalid_argument("cell_pressure_ref out of range");
  try {
    S355_HIP(hipStreamCreate(&stream));
    scratch_ = scratch_for_current_device();
    upload_mesh();
    allocate_fields();
    // the limiters' vertex arrays and the least-squares matrices, so that a
    // gradient allocates nothing
    if (par.limiter != 0) ensure_vertices();
    if (par.grad_scheme == 1) ensure_ls();
    S355_HIP(hipStreamSynchronize(stream));
  } catch (...) {
    free_all();
    throw;
  }
}

Context::~Context() { free_all(); }

Tiles Context::make_tiles(const std::vector<int>& case_row, int tile) {
  Tiles t;
  t.n = case_row.back();
  t.tile = tile;
  t.ncase = ncase;
  if (ncase == 1) {
    t.ntile = std::max(1, (t.n + tile - 1) / tile);
    return t;
  }
  std::vector<int> lo, tc, ct(ncase + 1);
  for (int c = 0; c < ncase; ++c) {
    ct[c] = static_cast<int>(lo.size());
    for (int r = case_row[c]; r < case_row[c + 1]; r += tile) {
      lo.push_back(r);
      tc.push_back(c);
    }
  }
  ct[ncase] = static_cast<int>(lo.size());
  t.ntile = static_cast<int>(lo.size());
  lo.push_back(t.n);
  t.lo = upload(allocs_, lo, stream);
  t.tcase = upload(allocs_, tc, stream);
  t.case_tile = upload(allocs_, ct, stream);
  t.case_row = upload(allocs_, case_row, stream);
  S355_HIP(hipStreamSynchronize(stream));  // host staging vectors go out of scope
  if (t.ntile < 1) throw std::invalid_argument("empty batch");
  return t;
}

void Context::free_all() {
  if (stream) hipStreamSynchronize(stream);
  graphs_.reset();
  kwork_.timer = kwork32_.timer = nullptr;
  timers_[0].reset();
  timers_[1].reset();
  for (hipEvent_t e : all_events_) hipEventDestroy(e);
  all_events_.clear();
  pool_.clear();
  open_.clear();
  done_.clear();
  for (void* p : allocs_) hipFree(p);
  allocs_.clear();
  if (host_pinned_) hipHostFree(host_pinned_);
  host_pinned_ = nullptr;
  if (hdr_) hipHostFree(hdr_);
  hdr_ = nullptr;
  if (stream) hipStreamDestroy(stream);
  stream = nullptr;
}

// CellFaceInfo (geometry.cpp:391-752) on the device, from the vertices and
// the topology of the host ingest (which skipped the metrics), then copied
// back into the host mesh for the host-side consumers (result files, the
// monitor, mesh queries).
void Context::compute_geometry() {
  Mesh& m = mesh;
  const size_t Nc = m.n_cell, Nf = m.n_face, Nb = m.n_bnd;
  dm.Nv = m.n_vert;
  dm.vert = upload(allocs_, m.vert, stream);
  dm.c_vert = upload(allocs_, m.c_vert, stream);
  // topology used by the geometry only: freed once it is done
  uint8_t *poly = nullptr, *warped = nullptr;
  int* fv = nullptr;
  S355_HIP(hipMalloc(&poly, std::max<size_t>(Nc, 1)));
  S355_HIP(hipMalloc(&warped, std::max<size_t>(Nf, 1)));
  S355_HIP(hipMalloc(&fv, std::max<size_t>(4 * Nf, 1) * sizeof(int)));
  std::vector<uint8_t> hw(Nf);
  try {
    S355_HIP(hipMemcpyAsync(poly, m.c_poly.data(), Nc, hipMemcpyHostToDevice, stream));
    S355_HIP(hipMemcpyAsync(fv, m.f_v.data(), 4 * Nf * sizeof(int), hipMemcpyHostToDevice, stream));
    GeomArgs g;
    g.Nc = m.n_cell;
    g.Nf = m.n_face;
    g.Nb = m.n_bnd;
    g.vert = dm.vert;
    g.c_vert = dm.c_vert;
    g.c_poly = poly;
    g.cf_ptr = dm.cf_ptr;
    g.cf_pat = dm.cf_pat;
    g.f_v = fv;
    g.f_c1 = dm.f_c1;
    g.f_c2 = dm.f_c2;
    g.c_vol = alloc<double>(Nc);
    g.c_x = alloc<double>(3 * Nc);
    g.f_x = alloc<double>(3 * Nf);
    g.f_n = alloc<double>(3 * Nf);
    g.f_xpac = alloc<double>(3 * Nf);
    g.f_xnac = alloc<double>(3 * Nf);
    g.f_area = alloc<double>(Nf);
    g.f_lambda = alloc<double>(Nf);
    g.f_rlencos = alloc<double>(Nf);
    g.b_dist = alloc<double>(std::max<size_t>(Nb, 1));
    g.warped = warped;
    launch_geometry(g, stream);
    dm.c_vol = g.c_vol;
    dm.c_x = g.c_x;
    dm.f_x = g.f_x;
    dm.f_n = g.f_n;
    dm.f_xpac = g.f_xpac;
    dm.f_xnac = g.f_xnac;
    dm.f_area = g.f_area;
    dm.f_lambda = g.f_lambda;
    dm.f_rlencos = g.f_rlencos;
    dm.b_dist = g.b_dist;
    auto down = [&](hvec<double>& h, const double* d, size_t n) {
      h.resize(n);
      if (n) S355_HIP(hipMemcpyAsync(h.data(), d, n * sizeof(double), hipMemcpyDeviceToHost, stream));
    };
    down(m.c_vol, g.c_vol, Nc);
    down(m.c_x, g.c_x, 3 * Nc);
    down(m.f_x, g.f_x, 3 * Nf);
    down(m.f_n, g.f_n, 3 * Nf);
    down(m.f_xpac, g.f_xpac, 3 * Nf);
    down(m.f_xnac, g.f_xnac, 3 * Nf);
    down(m.f_area, g.f_area, Nf);
    down(m.f_lambda, g.f_lambda, Nf);
    down(m.f_rlencos, g.f_rlencos, Nf);
    down(m.b_dist, g.b_dist, Nb);
    if (Nf) S355_HIP(hipMemcpyAsync(hw.data(), warped, Nf, hipMemcpyDeviceToHost, stream));
    S355_HIP(hipStreamSynchronize(stream));
  } catch (...) {
    hipStreamSynchronize(stream);
    hipFree(poly);
    hipFree(warped);
    hipFree(fv);
    throw;
  }
  S355_HIP(hipFree(poly));
  S355_HIP(hipFree(warped));
  S355_HIP(hipFree(fv));
  m.n_warped = 0;
  for (uint8_t w : hw) m.n_warped += w;
  for (size_t b = 0; b < Nb; ++b)
    if (m.b_dist[b] < 0.)
      throw std::runtime_error("mesh: wall distance is negative at boundary " + std::to_string(b));
}

void Context::upload_mesh() {
  const Mesh& m = mesh;
  dm.Nc = m.n_cell;
  dm.Nf = m.n_face;
  dm.Nb = m.n_bnd;
  dm.nnz = m.nnz();
  dm.f_c1 = upload(allocs_, m.f_c1, stream);
  dm.f_c2 = upload(allocs_, m.f_c2, stream);
  dm.cf_ptr = upload(allocs_, m.cf_ptr, stream);
  dm.cf_pat = upload(allocs_, m.cf_pat, stream);
  dm.rowptr = upload(allocs_, m.rowptr, stream);
  dm.col = upload(allocs_, m.col, stream);
  dm.diag = upload(allocs_, m.diag, stream);
  compute_geometry();
  std::vector<int> btype(m.n_bnd);
  for (int b = 0; b < m.n_bnd; ++b) {
    btype[b] = m.type1_of_rid(m.b_rid[b]);
    if (btype[b] == kOutlet) has_outlet_ = true;
  }
  dm.b_type = upload(allocs_, btype, stream);
  // wall-adjacent cells with their wall faces in ascending order
  std::vector<int> wcell, wptr(1, 0), wface;
  {
    std::vector<std::vector<int>> per(m.n_cell);
    for (int b = 0; b < m.n_bnd; ++b)
      if (btype[b] == kWall) per[m.f_c1[b]].push_back(b);
    for (int c = 0; c < m.n_cell; ++c) {
      if (per[c].empty()) continue;
      wcell.push_back(c);
      for (int f : per[c]) wface.push_back(f);
      wptr.push_back(static_cast<int>(wface.size()));
    }
  }
  // processing orders: cells by input element, boundary faces in place,
  // interior faces by the place of their first cell (then of their second)
  {
    std::vector<int> cord(m.n_cell), cpos(m.n_cell), ford(m.n_face);
    for (int c = 0; c < m.n_cell; ++c) cord[c] = c;
    std::sort(cord.begin(), cord.end(), [&](int a, int b) { return m.c_elem[a] < m.c_elem[b]; });
    for (int t = 0; t < m.n_cell; ++t) cpos[cord[t]] = t;
    for (int f = 0; f < m.n_face; ++f) ford[f] = f;
    std::sort(ford.begin() + m.n_bnd, ford.end(), [&](int a, int b) {
      const int a1 = cpos[m.f_c1[a]], b1 = cpos[m.f_c1[b]];
      return a1 < b1 || (a1 == b1 && cpos[m.f_c2[a]] < cpos[m.f_c2[b]]);
    });
    dm.c_order = upload(allocs_, cord, stream);
    dm.f_order = upload(allocs_, ford, stream);
    // the half-face streams in that order, packed into blocks of whole cells
    auto stream_of = [&](const hvec<int>& faces, const hvec<int>* slots, const int*& blk, int& nblk) {
      std::vector<int4> rec;
      std::vector<int> b(1, 0);
      rec.reserve(static_cast<size_t>(m.cf_ptr[m.n_cell]));
      int cells = 0;  // in the current block: at most one wave's worth
      for (int t = 0; t < m.n_cell; ++t) {
        const int i = cord[t], e0 = m.cf_ptr[i], e1 = m.cf_ptr[i + 1];
        if (e1 - e0 > kBlock) throw std::runtime_error("mesh: a cell with more faces than a block has lanes");
        if (static_cast<int>(rec.size()) - b.back() + (e1 - e0) > kBlock || cells == kWave) {
          b.push_back(static_cast<int>(rec.size()));
          cells = 0;
        }
        ++cells;
        for (int e = e0; e < e1; ++e) {
          const int f = faces[e];
          const int nb = f < m.n_bnd ? 0 : (m.f_c1[f] == i ? m.f_c2[f] : ~m.f_c1[f]);
          rec.push_back(make_int4(i, f, nb, slots ? (*slots)[e] : -1));
        }
      }
      b.push_back(static_cast<int>(rec.size()));
      nblk = static_cast<int>(b.size()) - 1;
      blk = upload(allocs_, b, stream);
      return upload(allocs_, rec, stream);
    };
    dm.hf = stream_of(m.cf_pat, &m.cf_slot, dm.hf_blk, dm.hf_nblk);
    S355_HIP(hipStreamSynchronize(stream));
    dm.hfs = stream_of(m.cf_sorted, nullptr, dm.hfs_blk, dm.hfs_nblk);
    S355_HIP(hipStreamSynchronize(stream));  // host staging vectors go out of scope
  }
  dm.ctile = make_tiles(lay.cell, kBlock);
  dm.btile = make_tiles(lay.bnd, kBlock);
  dm.n_wcell = static_cast<int>(wcell.size());
  dm.wc_cell = upload(allocs_, wcell, stream);
  dm.wc_ptr = upload(allocs_, wptr, stream);
  dm.wc_face = upload(allocs_, wface, stream);
  S355_HIP(hipStreamSynchronize(stream));  // host staging vectors go out of scope
}

void Context::allocate_fields() {
  const size_t Nc = dm.Nc, Nb = dm.Nb, Nf = dm.Nf, nnz = dm.nnz;
  double** cell1[] = {&df.R, &df.U, &df.V, &df.W, &df.P, &df.T, &df.TE, &df.ED, &df.dp, &df.mu, &df.mut, &df.apr,
                      &df.bu, &df.bv, &df.bw, &df.bp, &df.bs};
  for (double** p : cell1) *p = alloc<double>(Nc);
  double** cell3[] = {&df.dUdX, &df.dVdX, &df.dWdX, &df.dPdX, &df.dPhidX};
  for (double** p : cell3) *p = alloc<double>(3 * Nc);
  double** bnd1[] = {&df.BRo, &df.BU, &df.BV, &df.BW, &df.BPre, &df.BTE, &df.BED, &df.BTem};
  for (double** p : bnd1) *p = alloc<double>(Nb);
  df.shear = alloc<double>(3 * Nb);
  df.RUFace = alloc<double>(Nf);
  df.Au = alloc<double>(nnz);
  df.Ap = alloc<double>(nnz);
  df.As = alloc<double>(nnz);
  u0_ = alloc<double>(Nc);
  v0_ = alloc<double>(Nc);
  w0_ = alloc<double>(Nc);
  kw_.prod = alloc<double>(Nc);
  kw_.tes = alloc<double>(Nc);
  kw_.eds = alloc<double>(Nc);
  kw_.viste = alloc<double>(Nc);
  kw_.visted = alloc<double>(Nc);
  kw_.apte = alloc<double>(Nc);
  kw_.aped = alloc<double>(Nc);
  kw_.kcond = alloc<double>(Nc);
  kw_.zero = alloc<double>(Nc);
  // zero everything once so that get_field before any phase is defined
  double* all_cell[] = {df.R, df.U, df.V, df.W, df.P, df.T, df.TE, df.ED, df.dp, df.mu, df.mut, df.apr,
                        df.bu, df.bv, df.bw, df.bp, df.bs};
  for (double* p : all_cell) S355_HIP(hipMemsetAsync(p, 0, Nc * sizeof(double), stream));
  for (double* p : {df.dUdX, df.dVdX, df.dWdX, df.dPdX, df.dPhidX})
    S355_HIP(hipMemsetAsync(p, 0, 3 * Nc * sizeof(double), stream));
  for (double* p : {df.BRo, df.BU, df.BV, df.BW, df.BPre, df.BTE, df.BED, df.BTem})
    S355_HIP(hipMemsetAsync(p, 0, std::max<size_t>(Nb, 1) * sizeof(double), stream));
  S355_HIP(hipMemsetAsync(df.shear, 0, std::max<size_t>(3 * Nb, 1) * sizeof(double), stream));
  S355_HIP(hipMemsetAsync(df.RUFace, 0, Nf * sizeof(double), stream));
  for (double* p : {df.Au, df.Ap, df.As}) S355_HIP(hipMemsetAsync(p, 0, nnz * sizeof(double), stream));
  // scratch for reductions: per-tile partials of the Krylov kernels (up to
  // m + 1 columns) and of the cell / boundary reductions (up to 3 columns)
  max_restart_ = 4;
  for (int k = 0; k < 3; ++k) max_restart_ = std::max(max_restart_, std::min(par.ksp[k].restart, kMaxRestart));
  for (int r = 0; r < 5; ++r) kwork_.tiles[r] = make_tiles(lay.cell, kBlock << r);
  const size_t tiles1 = static_cast<size_t>(kwork_.tiles[0].ntile);
  partial_len_ = std::max<size_t>({tiles1 * (max_restart_ + 1), 3 * static_cast<size_t>(dm.btile.ntile),
                                   3 * static_cast<size_t>(dm.ctile.ntile)});
  partial_ = alloc<double>(partial_len_);
  scalars_ = alloc<double>(32);
  res_ = alloc<double>(10 * static_cast<size_t>(ncase));
  rate_ = alloc<double>(5 * static_cast<size_t>(ncase));
  flags_ = alloc<int>(8);
  S355_HIP(hipMemsetAsync(scalars_, 0, 32 * sizeof(double), stream));
  S355_HIP(hipMemsetAsync(res_, 0, 10 * sizeof(double) * ncase, stream));
  S355_HIP(hipMemsetAsync(rate_, 0, 5 * sizeof(double) * ncase, stream));
  S355_HIP(hipMemsetAsync(flags_, 0, 8 * sizeof(int), stream));
  // Krylov workspace
  kwork_.n = static_cast<int>(Nc);
  kwork_.ncase = ncase;
  kwork_.mmax = max_restart_;
  // basis columns padded to a multiple of 8 elements, so that every column
  // starts 16-byte aligned for the vector loads of the Gram-Schmidt kernels
  const size_t ldp = (Nc + 7) / 8 * 8;
  kwork_.ld = ldp;
  kwork_.V = alloc<double>((static_cast<size_t>(max_restart_) + 1) * ldp);
  kwork_.r = alloc<double>(Nc);
  kwork_.pcinv = alloc<double>(Nc);
  kwork_.partial = partial_;
  kwork_.partial_len = partial_len_;
  kwork_.st = alloc<KState>(ncase);
  S355_HIP(hipHostMalloc(&host_pinned_, 4096 + 10 * sizeof(double) * ncase, hipHostMallocDefault));
  S355_HIP(hipHostMalloc(reinterpret_cast<void**>(&hdr_), sizeof(KHeader) * ncase, hipHostMallocDefault));
  kwork_.host = hdr_;
}

void Context::set_params(const s355_params& p) {
  if (p.density_model != 0) throw std::invalid_argument("unsupported model switch");
  for (int k = 0; k < 3; ++k)
    if (std::min(p.ksp[k].restart, kMaxRestart) > max_restart_)
      throw std::invalid_argument("restart larger than the workspace built at context creation");
  int min_cells = dm.Nc;
  for (int k = 0; k < ncase; ++k) min_cells = std::min(min_cells, lay.cell[k + 1] - lay.cell[k]);
  if (p.cell_pressure_ref < 0 || p.cell_pressure_ref >= min_cells) throw std::invalid_argument("cell_pressure_ref");
  if (ncase > 1)
    for (int k = 0; k < 3; ++k)
      if (p.ksp[k].type == S355_KSP_CG) throw std::invalid_argument("a batch of cases solves every system with GMRES");
  if (p.inner_order < 0 || p.inner_order > 2) throw std::invalid_argument("inner_order: 0, 1 (RCM) or 2 (input)");
  if (have_order_ && p.inner_order != 0 && p.inner_order != order_kind_)
    throw std::invalid_argument("inner_order: the other inner order was built already");
  if (p.grad_scheme < 0 || p.grad_scheme > 1) throw std::invalid_argument("grad_scheme: 0 Green-Gauss, 1 least squares");
  if (p.assembly_fp32 != 0 && p.assembly_fp32 != 1) throw std::invalid_argument("assembly_fp32: 0 or 1");
  par = p;
  if (par.limiter != 0) ensure_vertices();
  if (par.grad_scheme == 1) ensure_ls();
  // captured phases hold the old parameters in their kernel arguments
  phase_graphs_.reset();
  phase_seen_ = 0;
}

FieldRef Context::field(int id) {
  const size_t Nc = dm.Nc, Nb = dm.Nb, Nf = dm.Nf, nnz = dm.nnz;
  switch (id) {
    case S355_F_R: return {df.R, Nc};
    case S355_F_U: return {df.U, Nc};
    case S355_F_V: return {df.V, Nc};
    case S355_F_W: return {df.W, Nc};
    case S355_F_P: return {df.P, Nc};
    case S355_F_T: return {df.T, Nc};
    case S355_F_TE: return {df.TE, Nc};
    case S355_F_ED: return {df.ED, Nc};
    case S355_F_DELTAP: return {df.dp, Nc};
    case S355_F_VISLAM: return {df.mu, Nc};
    case S355_F_VISTUR: return {df.mut, Nc};
    case S355_F_APR: return {df.apr, Nc};
    case S355_F_DUDX: return {df.dUdX, 3 * Nc};
    case S355_F_DVDX: return {df.dVdX, 3 * Nc};
    case S355_F_DWDX: return {df.dWdX, 3 * Nc};
    case S355_F_DPDX: return {df.dPdX, 3 * Nc};
    case S355_F_DPHIDX: return {df.dPhidX, 3 * Nc};
    case S355_F_BRO: return {df.BRo, Nb};
    case S355_F_BU: return {df.BU, Nb};
    case S355_F_BV: return {df.BV, Nb};
    case S355_F_BW: return {df.BW, Nb};
    case S355_F_BPRE: return {df.BPre, Nb};
    case S355_F_BTE: return {df.BTE, Nb};
    case S355_F_BED: return {df.BED, Nb};
    case S355_F_SHEAR: return {df.shear, 3 * Nb};
    case S355_F_RUFACE: return {df.RUFace, Nf};
    case S355_F_AU: return {df.Au, nnz};
    case S355_F_AP: return {df.Ap, nnz};
    case S355_F_AS: return {df.As, nnz};
    case S355_F_BU_RHS: return {df.bu, Nc};
    case S355_F_BV_RHS: return {df.bv, Nc};
    case S355_F_BW_RHS: return {df.bw, Nc};
    case S355_F_BP_RHS: return {df.bp, Nc};
    case S355_F_BS_RHS: return {df.bs, Nc};
    case S355_F_BTEM: return {df.BTem, Nb};
    case S355_F_UNP: case S355_F_VNP: case S355_F_WNP: case S355_F_TNP: case S355_F_TEP: case S355_F_EDP:
    case S355_F_UNP2: case S355_F_VNP2: case S355_F_WNP2: case S355_F_TNP2: case S355_F_TEP2: case S355_F_EDP2: {
      double* old[12] = {df.Unp, df.Vnp, df.Wnp, df.Tnp, df.TEp, df.EDp,
                         df.Unp2, df.Vnp2, df.Wnp2, df.Tnp2, df.TEp2, df.EDp2};
      double* p = old[id - S355_F_UNP];
      if (!p) throw std::invalid_argument("old time levels exist once an unsteady run has saved them");
      return {p, Nc};
    }
    case S355_F_KE_PROD: return {kw_.prod, Nc};
    case S355_F_KE_TESRC: return {kw_.tes, Nc};
    case S355_F_KE_EDSRC: return {kw_.eds, Nc};
    case S355_F_KE_APTE: return {kw_.apte, Nc};
    case S355_F_KE_APED: return {kw_.aped, Nc};
    case S355_F_KE_VISTE: return {kw_.viste, Nc};
    case S355_F_KE_VISED: return {kw_.visted, Nc};
    default: throw std::invalid_argument("unknown field id " + std::to_string(id));
  }
}

void Context::sync() { S355_HIP(hipStreamSynchronize(stream)); }

// InitFlowField (navier.cpp:588-697), evaluated on the host once and uploaded.
void Context::init_flow_field() {
  const int Nc = dm.Nc, Nb = dm.Nb;
  std::vector<double> U(Nc), V(Nc), W(Nc), R(Nc), T(Nc), P(Nc, 0.0), mu(Nc), mut(Nc, 0.0), k(Nc, 0.0), e(Nc, 0.0);
  auto find = [&](int rid) -> const s355_region* {
    for (const auto& r : mesh.regions)
      if (r.rid == rid) return &r;
    return nullptr;
  };
  auto apply = [&](int i, const s355_region& r) {
    U[i] = r.init[0];
    V[i] = r.init[1];
    W[i] = r.init[2];
    R[i] = r.init[3];
    T[i] = r.init[4];
    P[i] = 0.;
    mu[i] = r.init[5];
    mut[i] = 0.;
    if (par.turb_model == 1) {
      k[i] = r.init[6];
      e[i] = r.init[7];
      mut[i] = R[i] * kCmuInit * k[i] * k[i] / (e[i] + kSmallInit);
    }
  };
  // region 0 (the default initial state of a case) first, then each cell's
  // own region; in a batch the region ids of case k start at k * rid_stride
  for (int k = 0; k < ncase; ++k)
    if (const s355_region* r0 = find(k * lay.rid_stride))
      for (int i = lay.cell[k]; i < lay.cell[k + 1]; ++i) apply(i, *r0);
  for (int i = 0; i < Nc; ++i) {
    const s355_region* r = find(mesh.c_rid[i]);
    if (!r) throw std::invalid_argument("cell region missing");
    apply(i, *r);
  }
  std::vector<double> BU(Nb, 0.0), BV(Nb, 0.0), BW(Nb, 0.0), BRo(Nb, 0.0), BTE(Nb, 0.0), BED(Nb, 0.0),
      BTem(Nb, 0.0);
  for (int b = 0; b < Nb; ++b) {
    const s355_region* r = find(mesh.b_rid[b]);
    if (r->type1 == kWall && r->type2 == 0) BTem[b] = r->fixed_value;  // isothermal wall
    if (r->type1 == kInlet) {
      BU[b] = r->init[0];
      BV[b] = r->init[1];
      BW[b] = r->init[2];
      BRo[b] = r->init[4];
      BTem[b] = r->init[5];
      BTE[b] = r->init[6];
      BED[b] = r->init[7];
    }
  }
  auto put = [&](double* d, const std::vector<double>& h) {
    if (!h.empty()) S355_HIP(hipMemcpyAsync(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, stream));
  };
  put(df.U, U);
  put(df.V, V);
  put(df.W, W);
  put(df.R, R);
  put(df.T, T);
  put(df.P, P);
  put(df.mu, mu);
  put(df.mut, mut);
  put(df.TE, k);
  put(df.ED, e);
  S355_HIP(hipMemsetAsync(df.dp, 0, Nc * sizeof(double), stream));
  put(df.BU, BU);
  put(df.BV, BV);
  put(df.BW, BW);
  put(df.BRo, BRo);
  put(df.BTE, BTE);
  put(df.BED, BED);
  put(df.BTem, BTem);
  S355_HIP(hipMemsetAsync(df.RUFace, 0, dm.Nf * sizeof(double), stream));
  sync();  // host staging vectors go out of scope
}

void Context::bc_velocity() { launch_bc_velocity(dm, df, has_outlet_, partial_, rate_, stream); }
void Context::bc_pressure() { launch_bc_copy(dm, df.P, df.BPre, stream); }
void Context::bc_deltap() { launch_bc_copy(dm, df.dp, df.BPre, stream); }

// Green-Gauss gradients of nf fields, then the limiter on each (Gradient,
// navier_gradient.cpp:12-74)
void Context::grad(int nf, const double* const* phi, const double* const* bphi, double* const* out) {
  if (par.grad_scheme == 1) {
    ensure_ls();
    launch_gradient_ls(dm, nf, phi, bphi, out, ls_inv_, stream);
  } else {
    launch_gradient(dm, nf, phi, bphi, out, stream);
  }
  if (par.limiter == 0) return;
  ensure_vertices();
  for (int k = 0; k < nf; ++k)
    launch_limiter(dm, par.limiter, phi[k], out[k], lim_vmax_, lim_vmin_, lim_tmp_, stream);
}

void Context::gradient(int field_id, int bfield_id, int out_id) {
  FieldRef f = field(field_id), b = field(bfield_id), g = field(out_id);
  if (f.n != static_cast<size_t>(dm.Nc) || b.n != static_cast<size_t>(dm.Nb) || g.n != 3 * static_cast<size_t>(dm.Nc))
    throw std::invalid_argument("gradient: expects a cell field, a boundary field and a cell 3-vector field");
  const double* phi[1] = {f.p};
  const double* bphi[1] = {b.p};
  double* out[1] = {g.p};
  grad(1, phi, bphi, out);
}

// The same gradient with the limiters' scratch (MLP: the extrema at every
// vertex, 2 Nvert doubles; WENO: the limited copy, 3 Ncel doubles) in the
// fixed scratch array: no allocation.  Both are written in full before they
// are read, so the array needs no clearing; calls that overlap on one GPU
// share the array and race.
void Context::gradient_noalloc(int field_id, int bfield_id, int out_id) {
  FieldRef f = field(field_id), b = field(bfield_id), g = field(out_id);
  if (f.n != static_cast<size_t>(dm.Nc) || b.n != static_cast<size_t>(dm.Nb) || g.n != 3 * static_cast<size_t>(dm.Nc))
    throw std::invalid_argument("gradient: expects a cell field, a boundary field and a cell 3-vector field");
  size_t need = 0;
  if (par.limiter == 2) need = 2 * static_cast<size_t>(dm.Nv);
  if (par.limiter == 3) need = 3 * static_cast<size_t>(dm.Nc);
  if (need > kScratchDoubles)
    throw std::invalid_argument("gradient_noalloc: the limiter needs " + std::to_string(need) +
                                " doubles of scratch, above the fixed " + std::to_string(kScratchDoubles) +
                                " (MLP: 2 x vertices, WENO: 3 x cells)");
  if (par.limiter != 0 && !dm.c_vert) throw std::logic_error("gradient_noalloc: limiter vertex arrays missing");
  if (!scratch_) throw std::logic_error("gradient_noalloc: no scratch address");
  const double* phi[1] = {f.p};
  const double* bphi[1] = {b.p};
  double* out[1] = {g.p};
  if (par.grad_scheme == 1) {
    if (!ls_inv_) throw std::logic_error("gradient_noalloc: least-squares matrices missing");
    launch_gradient_ls(dm, 1, phi, bphi, out, ls_inv_, stream);
  } else {
    launch_gradient(dm, 1, phi, bphi, out, stream);
  }
  if (par.limiter == 0) return;
  double* vmax = scratch_;
  double* vmin = scratch_ + dm.Nv;
  launch_limiter(dm, par.limiter, f.p, g.p, vmax, vmin, scratch_, stream);
}

void Context::build_velocity_matrix() {
  bc_velocity();
  bc_pressure();
  const double* phi[4] = {df.U, df.V, df.W, df.P};
  const double* bphi[4] = {df.BU, df.BV, df.BW, df.BPre};
  double* out[4] = {df.dUdX, df.dVdX, df.dWdX, df.dPdX};
  grad(4, phi, bphi, out);
  const int scheme = par.steady ? 0 : par.time_scheme;
  if (scheme) ensure_transient();
  launch_build_momentum(dm, df, par.urf[0], par.gravity, flags_, scheme, par.dt, fp32_assembly(), stream);
}

// the inverse least-squares normal matrices, built once
void Context::ensure_ls() {
  if (ls_inv_) return;
  ls_inv_ = alloc<double>(6 * static_cast<size_t>(dm.Nc));
  launch_ls_setup(dm, ls_inv_, stream);
}

// vertex arrays of the limiters: each cell's 8 vertices, the coordinates, and
// the cells around each vertex (each cell once, ascending)
void Context::ensure_vertices() {
  if (dm.v2c) return;
  const int nv = mesh.n_vert, nc = mesh.n_cell;
  std::vector<int> ptr(nv + 1, 0);
  auto distinct = [&](int i, int j) {
    const int* h = &mesh.c_vert[8 * static_cast<size_t>(i)];
    for (int q = 0; q < j; ++q)
      if (h[q] == h[j]) return false;
    return true;
  };
  for (int i = 0; i < nc; ++i)
    for (int j = 0; j < 8; ++j)
      if (distinct(i, j)) ++ptr[mesh.c_vert[8 * static_cast<size_t>(i) + j] + 1];
  for (int v = 0; v < nv; ++v) ptr[v + 1] += ptr[v];
  std::vector<int> v2c(ptr[nv]), fill(ptr.begin(), ptr.end() - 1);
  for (int i = 0; i < nc; ++i)
    for (int j = 0; j < 8; ++j)
      if (distinct(i, j)) v2c[fill[mesh.c_vert[8 * static_cast<size_t>(i) + j]]++] = i;
  dm.v2c_ptr = upload(allocs_, ptr, stream);
  dm.v2c = upload(allocs_, v2c, stream);
  lim_vmax_ = alloc<double>(std::max(nv, 1));
  lim_vmin_ = alloc<double>(std::max(nv, 1));
  lim_tmp_ = alloc<double>(3 * static_cast<size_t>(dm.Nc));
  sync();  // host staging vectors go out of scope
}

// old time levels and the energy time-term arrays, zeroed once
void Context::ensure_transient() {
  if (df.Unp) return;
  const size_t Nc = dm.Nc;
  double** old[] = {&df.Unp, &df.Vnp, &df.Wnp, &df.Tnp, &df.TEp, &df.EDp,
                    &df.Unp2, &df.Vnp2, &df.Wnp2, &df.Tnp2, &df.TEp2, &df.EDp2, &kw_.esrc, &kw_.ape};
  for (double** p : old) {
    *p = alloc<double>(Nc);
    fill_zero(*p, Nc * sizeof(double), stream);
  }
}

void Context::save_old_data() {
  if (par.steady) throw std::logic_error("SaveTransientOldData: the run is steady");
  ensure_transient();
  launch_save_old(dm, df, par.time_scheme, par.turb_model == 1, stream);
}

void Context::unsteady_sources(int isca) {
  if (isca != 1 && isca != 2) throw std::invalid_argument("unsteady_sources: isca 1 (T) or 2 (k and epsilon)");
  if (par.steady) return;
  ensure_transient();
  launch_unsteady_scalar(dm, df, kw_, isca, par.time_scheme, par.dt, stream);
}

namespace {
void fill_info(const KResult& r, s355_solve_info* info) {
  if (!info) return;
  info->iters = r.iters;
  info->reason = r.reason;
  info->rnorm = r.rnorm;
  info->rnorm0 = r.rnorm0;
  info->refinements = r.refinements;
  info->pad = 0;
}
}  // namespace

void Context::ensure_ilu() {
  if (have_levels_) return;
  std::vector<int> lrows, urows;
  ilu_levels(dm.Nc, mesh.rowptr.data(), mesh.col.data(), mesh.diag.data(), lrows, levels_.lptr, urows, levels_.uptr);
  levels_.lrows = upload(allocs_, lrows, stream);
  levels_.urows = upload(allocs_, urows, stream);
  // positions of each row in the two level orders (the sweeps' z spaces)
  std::vector<int> lpos(dm.Nc), upos(dm.Nc), l2u(dm.Nc);
  for (int p = 0; p < dm.Nc; ++p) {
    lpos[lrows[p]] = p;
    upos[urows[p]] = p;
  }
  for (int p = 0; p < dm.Nc; ++p) l2u[p] = upos[lrows[p]];
  levels_.lpos = upload(allocs_, lpos, stream);
  {
    std::vector<int> sptr, sp, sd, su, sr, sq;
    ilu_steps(dm.Nc, mesh.rowptr.data(), mesh.col.data(), mesh.diag.data(), lrows, sptr, sp, sd, su, sr, sq);
    levels_.steps.ptr = upload(allocs_, sptr, stream);
    levels_.steps.p = upload(allocs_, sp, stream);
    levels_.steps.d = upload(allocs_, sd, stream);
    levels_.steps.u = upload(allocs_, su, stream);
    levels_.steps.r = upload(allocs_, sr, stream);
    levels_.steps.q = upload(allocs_, sq, stream);
    sync();  // host staging vectors go out of scope
  }
  levels_.l2u = upload(allocs_, l2u, stream);
  levels_.upos = upload(allocs_, upos, stream);
  // level-ELL copies of the factors for the sweeps
  for (int up = 0; up < 2; ++up) {
    IluEll& e = up ? levels_.uell : levels_.lell;
    std::vector<int> ecol, esrc, len, dsrc;
    ilu_ell(dm.Nc, mesh.rowptr.data(), mesh.col.data(), mesh.diag.data(), up ? urows : lrows,
            up ? levels_.uptr : levels_.lptr, up ? upos : lpos, up != 0, e.off, ecol, esrc, len, dsrc);
    e.slots = esrc.size();
    e.col = upload(allocs_, ecol, stream);
    e.src = upload(allocs_, esrc, stream);
    e.len = upload(allocs_, len, stream);
    if (up) e.dsrc = upload(allocs_, dsrc, stream);
    sync();  // host staging vectors go out of scope
  }
  kwork_.lu = alloc<double>(dm.nnz);
  kwork_.ilu_lval = alloc<double>(levels_.lell.slots);
  kwork_.ilu_uval = alloc<double>(levels_.uell.slots);
  kwork_.ilu_dval = alloc<double>(dm.Nc);
  kwork_.ilu_zl = alloc<double>(dm.Nc);
  kwork_.ilu_zu = alloc<double>(dm.Nc);
  kwork_.ilu = &levels_;
  sync();  // host staging of the level lists
  have_levels_ = true;
}

void Context::ensure_mixed() {
  if (a32_) return;
  const size_t Nc = dm.Nc;
  a32_ = alloc<float>(dm.nnz);
  kwork32_.r = alloc<float>(Nc);
  kwork32_.lu = alloc<float>(dm.nnz);
  kwork32_.n = dm.Nc;
  kwork32_.mmax = max_restart_;
  kwork32_.ld = kwork_.ld;
  kwork32_.V = alloc<float>((static_cast<size_t>(max_restart_) + 1) * kwork32_.ld);
  kwork32_.pcinv = alloc<float>(Nc);
  kwork32_.partial = partial_;
  kwork32_.partial_len = partial_len_;
  kwork32_.st = kwork_.st;
  kwork32_.host = kwork_.host;
  kwork32_.ncase = kwork_.ncase;
  for (int r = 0; r < 5; ++r) kwork32_.tiles[r] = kwork_.tiles[r];
  ir_.r64 = alloc<double>(Nc);
  ir_.r32 = alloc<float>(Nc);
  ir_.e32 = alloc<float>(Nc);
  ir_.inv32 = alloc<float>(Nc);
  ir_.partial = partial_;
  ir_.partial_len = partial_len_;
  ir_.scal = scalars_ + 20;
  ir_.host = reinterpret_cast<double*>(static_cast<char*>(host_pinned_) + 3584);
}

int Context::run_solve(int which, const double* b, double* x, const s355_ksp_opts& o, bool zero_guess,
                       s355_solve_info* info, int slot) {
  const double* val = which == 0 ? df.Au : (which == 1 ? df.Ap : df.As);
  KResult r;
  if (par.graphs && !graphs_) graphs_ = std::make_unique<GraphCache>();
  kwork_.graphs = par.graphs ? graphs_.get() : nullptr;
  kwork32_.graphs = kwork_.graphs;
  kwork_.per_case = &case_info[slot];
  kwork32_.per_case = &case_info[slot];
  case_info[slot].assign(ncase, KResult{});
  if (o.pc == S355_PC_ILU0) ensure_ilu();
  if (par.precision == S355_MIXED) {
    // FP32 inner solves on the rounded matrix inside FP64 refinement
    ensure_mixed();
    kwork32_.ilu = kwork_.ilu;
    if (kwork32_.ilu && !kwork32_.ilu_lval) {
      kwork32_.ilu_lval = alloc<float>(levels_.lell.slots);
      kwork32_.ilu_uval = alloc<float>(levels_.uell.slots);
      kwork32_.ilu_dval = alloc<float>(dm.Nc);
      kwork32_.ilu_zl = alloc<float>(dm.Nc);
      kwork32_.ilu_zu = alloc<float>(dm.Nc);
    }
    CsrView<float> A{dm.Nc, dm.rowptr, dm.col, dm.diag, a32_, dm.nnz};
    CsrView<double> A64{dm.Nc, dm.rowptr, dm.col, dm.diag, val, dm.nnz};
    const bool tune = par.autotune && o.type == S355_KSP_GMRES && !kwork32_.tuned;
    const bool ordered = use_order(o);
    if (!ordered) convert_to_float(val, a32_, dm.nnz, stream);  // the inner order fills its own copy
    if (ordered) {
      // the Arnoldi cycle on the rows in the inner order, tuned there
      fill_order(val, o.pc == S355_PC_JACOBI);
      CsrView<float> Ao{dm.Nc, ord_rowptr_, ord_col_, ord_diag_, a32o_, dm.nnz};
      if (tune) tune_ordered(Ao, std::min(o.restart, kMaxRestart));
      KWork w = ordered_work(false);
      r = ir_solve(A64, Ao, b, x, o, zero_guess, par.ir_max, par.ir_inner_rtol, w, ir_, stream);
    } else {
      if (o.type == S355_KSP_GMRES && (tune || kwork32_.lanes == 0)) prepare_sell(val, true, tune);
      if (tune) autotune<float>(A, kwork32_, std::min(o.restart, kMaxRestart), stream);
      r = ir_solve(A64, A, b, x, o, zero_guess, par.ir_max, par.ir_inner_rtol, kwork32_, ir_, stream);
    }
  } else {
    CsrView<double> A{dm.Nc, dm.rowptr, dm.col, dm.diag, val, dm.nnz};
    const bool tune = par.autotune && o.type == S355_KSP_GMRES && !kwork_.tuned;
    if (o.type == S355_KSP_GMRES && (tune || kwork_.lanes == 0)) prepare_sell(val, false, tune);
    if (tune) autotune<double>(A, kwork_, std::min(o.restart, kMaxRestart), stream);
    r = krylov_solve<double>(A, b, x, o, zero_guess, kwork_, stream);
  }
  if (ncase == 1) case_info[slot].assign(1, r);
  fill_info(r, info);
  return r.reason < 0 ? S355_NOT_CONVERGED : S355_OK;
}

int Context::solve_velocity(double rtol, int maxit, s355_solve_info info[3]) {
  s355_ksp_opts o = par.ksp[S355_KSP_VELOCITY];
  o.rtol = rtol;
  o.maxit = maxit;
  int rc = S355_OK;
  // The reference stops at the first failing component (the throw in
  // MPIStructure.cpp:279-283 leaves V and W unsolved).
  const double* b[3] = {df.bu, df.bv, df.bw};
  double* x[3] = {df.U, df.V, df.W};
  for (int k = 0; k < 3; ++k) {
    rc = run_solve(0, b[k], x[k], o, false, &info[k], k);
    if (rc != S355_OK) {
      for (int j = k + 1; j < 3; ++j) info[j] = s355_solve_info{0, 0, 0.0, 0.0, 0, 0};
      return rc;
    }
  }
  return rc;
}

void Context::cal_ru_face() { launch_ruface(dm, df, stream); }
void Context::build_pressure_matrix() { launch_build_pressure(dm, df, fp32_assembly(), stream); }

int Context::solve_pressure(double rtol, int maxit, s355_solve_info* info) {
  s355_ksp_opts o = par.ksp[S355_KSP_PRESSURE];
  o.rtol = rtol;
  o.maxit = maxit;
  return run_solve(1, df.bp, df.dp, o, true, info, 3);
}

void Context::correct_fields() {
  bc_deltap();
  const double* phi[1] = {df.dp};
  const double* bphi[1] = {df.BPre};
  double* out[1] = {df.dPdX};
  grad(1, phi, bphi, out);
  launch_correct_fields(dm, df, par.cell_pressure_ref, par.urf[3], partial_, stream);
  launch_reduce_tiles(partial_, dm.ctile, 1, res_ + 3, 10, true, stream);
}

void Context::correct_ru_face() { launch_correct_ruface(dm, df, par.urf[3], stream); }

void Context::save_velocity() { launch_copy3(dm.Nc, df.U, df.V, df.W, u0_, v0_, w0_, stream); }

// The phases between the solves of a SIMPLE pass (navier.cpp:60-101).  Each
// is the sequence of member calls the host driver makes there; captured, it
// replays the same kernels with the same arguments, so both give the same
// bits.  No phase allocates after its first call or reads anything back.
void Context::phase_body(int phase) {
  switch (phase) {
    case S355_PHASE_MOMENTUM:
      build_velocity_matrix();
      save_velocity();
      break;
    case S355_PHASE_VELOCITY_RESIDUAL:
      velocity_residual();
      break;
    case S355_PHASE_PRESSURE:
      bc_velocity();
      cal_ru_face();
      build_pressure_matrix();
      break;
    case S355_PHASE_CORRECT:
      correct_fields();
      bc_velocity();
      correct_ru_face();
      break;
    case S355_PHASE_KEPS_K:
      kepsilon_sources();
      bc_kepsilon();
      unsteady_sources(2);  // nothing in steady runs
      build_scalar_matrix(2);
      break;
    case S355_PHASE_KEPS_EPS:
      kepsilon_post_k();
      build_scalar_matrix(3);
      eps_wall_rows();
      break;
    case S355_PHASE_KEPS_POST:
      kepsilon_post_eps();
      break;
    default:
      throw std::invalid_argument("run_phase: unknown phase");
  }
}

void Context::run_phase(int phase) {
  if (phase < 0 || phase >= S355_PHASE_COUNT) throw std::invalid_argument("run_phase: unknown phase");
  if (!par.graphs || !(phase_seen_ >> phase & 1u)) {
    phase_body(phase);
    phase_seen_ |= 1u << phase;
    return;
  }
  if (!phase_graphs_) phase_graphs_ = std::make_unique<GraphCache>();
  hipGraphExec_t g = phase_graphs_->get({static_cast<uintptr_t>(phase)}, stream, [&] { phase_body(phase); });
  S355_HIP(hipGraphLaunch(g, stream));
}

void Context::velocity_residual() {
  launch_residual3(dm, u0_, v0_, w0_, df.U, df.V, df.W, partial_, stream);
  launch_reduce_tiles(partial_, dm.ctile, 3, res_, 10, true, stream);
}

void Context::kepsilon_sources() { launch_keps_sources(dm, df, kw_, stream); }

void Context::bc_kepsilon() {
  launch_keps_wall(dm, df, kw_, std::pow(0.09, 0.25), std::pow(0.09, 0.75), stream);
  launch_keps_copy_bc(dm, df, stream);
}

void Context::bc_temperature() { launch_bc_temperature(dm, df, flags_, stream); }

void Context::energy_coefficients() { launch_energy_coef(dm, df, kw_, par.cp, par.prl, par.prte, stream); }

void Context::build_scalar_matrix(int isca) {
  double *phi, *bphi;
  const double *diff, *src, *app;
  if (isca == 1) {
    phi = df.T;
    bphi = df.BTem;
    diff = kw_.kcond;
    // unsteady: the time term written by unsteady_sources(1)
    if (!par.steady) ensure_transient();
    src = par.steady ? kw_.zero : kw_.esrc;
    app = par.steady ? kw_.zero : kw_.ape;
  } else if (isca == 2) {
    phi = df.TE;
    bphi = df.BTE;
    diff = kw_.viste;
    src = kw_.tes;
    app = kw_.apte;
  } else if (isca == 3) {
    phi = df.ED;
    bphi = df.BED;
    diff = kw_.visted;
    src = kw_.eds;
    app = kw_.aped;
  } else {
    throw std::invalid_argument("build_scalar_matrix: isca 1 (T), 2 (k) or 3 (epsilon)");
  }
  const double* p1[1] = {phi};
  const double* b1[1] = {bphi};
  double* g1[1] = {df.dPhidX};
  grad(1, p1, b1, g1);
  launch_build_scalar(dm, df, phi, bphi, diff, src, app, par.urf[5], flags_, fp32_assembly(), stream);
}

void Context::eps_wall_rows() { launch_dirichlet_rows(dm, df, stream); }

int Context::solve_scalar(int isca, double rtol, int maxit, s355_solve_info* info) {
  if (isca < 1 || isca > 3) throw std::invalid_argument("solve_scalar: isca 1 (T), 2 (k) or 3 (epsilon)");
  s355_ksp_opts o = par.ksp[S355_KSP_SCALAR];
  o.rtol = rtol;
  o.maxit = maxit;
  double* x = isca == 1 ? df.T : (isca == 2 ? df.TE : df.ED);
  return run_solve(2, df.bs, x, o, false, info, isca == 1 ? 6 : (isca == 2 ? 4 : 5));
}

void Context::kepsilon_post_k() {
  launch_clip(dm.Nc, df.TE, 1.e-9, stream);
  launch_eps_sources(dm, df, kw_, stream);
}

void Context::kepsilon_post_eps() {
  launch_clip(dm.Nc, df.ED, 1.e-12, stream);
  launch_turb_viscosity(dm, df, stream);
}

void Context::zero_residuals() { fill_zero(res_, 10 * sizeof(double) * ncase, stream); }

void Context::read_case_residuals(double* out) {
  // staged after the 4 KiB of fixed pinned slots
  double* h = reinterpret_cast<double*>(static_cast<char*>(host_pinned_) + 4096);
  S355_HIP(hipMemcpyAsync(h, res_, 10 * sizeof(double) * ncase, hipMemcpyDeviceToHost, stream));
  sync();
  std::memcpy(out, h, 10 * sizeof(double) * ncase);
}

void Context::read_residuals(double out[10]) {
  std::vector<double> all(10 * static_cast<size_t>(ncase));
  read_case_residuals(all.data());
  for (int j = 0; j < 10; ++j) {
    double v = all[j];
    for (int k = 1; k < ncase; ++k) v = std::max(v, all[10 * k + j]);
    out[j] = v;
  }
}

void Context::check_flags() {
  int* h = reinterpret_cast<int*>(static_cast<char*>(host_pinned_) + 3072);
  S355_HIP(hipMemcpyAsync(h, flags_, 2 * sizeof(int), hipMemcpyDeviceToHost, stream));
  sync();
  if (h[0]) throw std::runtime_error("unknown boundary type met in an assembly kernel");
  if (h[1]) throw std::runtime_error("reverse flow get out of inlet (scalar transport)");
}

// The case of each 64-row slice of a SELL pattern whose rows (slot rows
// srow, which are row ids in the pattern's own numbering) all belong to one
// case of a batch, else -1.
std::vector<int> Context::slice_cases(const std::vector<int>& ptr, const std::vector<int>& srow) const {
  const std::vector<int>& cr = lay.cell;
  auto case_of = [&](int r) { return static_cast<int>(std::upper_bound(cr.begin(), cr.end(), r) - cr.begin()) - 1; };
  const int nslice = static_cast<int>(ptr.size()) - 1;
  std::vector<int> scase(nslice, -1);
  for (int sl = 0; sl < nslice; ++sl) {
    const int p0 = sl * 64, p1 = std::min(dm.Nc, p0 + 64);
    const int c0 = case_of(srow[p0]);
    bool one = true;
    for (int p = p0 + 1; p < p1 && one; ++p) one = case_of(srow[p]) == c0;
    if (one) scase[sl] = c0;
  }
  return scase;
}

void Context::ensure_sell() {
  if (have_sell_) return;
  for (int v = 0; v < kSellVariants; ++v) {
    std::vector<int> ptr, row, col, src;
    sell_pattern(dm.Nc, mesh.rowptr.data(), mesh.col.data(), kSellSigma[v], ptr, row, col, src);
    SellPat pt;
    pt.slots = src.size();
    pt.ptr = upload(allocs_, ptr, stream);
    pt.row = upload(allocs_, row, stream);
    pt.col = upload(allocs_, col, stream);
    pt.src = upload(allocs_, src, stream);
    std::vector<int> scase;
    if (ncase > 1) {
      scase = slice_cases(ptr, row);
      pt.scase = upload(allocs_, scase, stream);
    }
    sync();  // host staging vectors go out of scope
    kwork_.sell[v] = kwork32_.sell[v] = pt;
  }
  have_sell_ = true;
}

bool Context::use_order(const s355_ksp_opts& o) {
  if (!par.inner_order || par.precision != S355_MIXED || o.type != S355_KSP_GMRES || o.pc == S355_PC_ILU0)
    return false;
  ensure_order();
  return order_ok_;
}

// The rows in the inner order -- 1: reverse Cuthill-McKee order of the
// pattern; 2: the order of the cells' elements in the input stream, before
// the partition sort (mesh generators write cells in a spatially coherent
// order, which the sort of equal keys scatters) -- then within windows of
// kSellSigma[1] positions by decreasing row length (stable), as a CSR by
// position (columns as positions, each row's entries by ascending column)
// and its SELL-64 copies with the columns as 16-bit offsets from the row,
// built once.  The windows are the SELL copies' own, so their slot rows are
// the positions themselves (no row map) except for the widest window.
// Ascending columns make the j-th entries of neighbouring rows neighbours
// too, so a wave's gathers of x for one slot fall in few cache lines; each
// row's sum runs in that order (the inner cycle is FP32 work inside FP64
// refinement, checked against the FP64 residual).
void Context::ensure_order() {
  if (have_order_) return;
  have_order_ = true;
  order_kind_ = par.inner_order;
  const int n = dm.Nc;
  std::vector<int> row;
  if (par.inner_order == 2) {
    // a batch's cases stay in their own position ranges: the element
    // indices of case k follow those of case k - 1
    row.resize(n);
    for (int p = 0; p < n; ++p) row[p] = p;
    std::sort(row.begin(), row.end(), [&](int a, int b) { return mesh.c_elem[a] < mesh.c_elem[b]; });
  } else {
    if (ncase > 1) return;  // the reverse Cuthill-McKee order is built for one case
    row = rcm_order(n, mesh.rowptr.data(), mesh.col.data());
  }
  auto len = [&](int r) { return mesh.rowptr[r + 1] - mesh.rowptr[r]; };
  // windows start at each case's first row, so a case's windows (and so its
  // rows' places) are the same in a batch as alone
  const int window = kSellSigma[1];
  for (int k = 0; k < ncase; ++k)
    for (int w0 = lay.cell[k]; w0 < lay.cell[k + 1]; w0 += window)
      std::stable_sort(row.begin() + w0, row.begin() + std::min(lay.cell[k + 1], w0 + window),
                       [&](int a, int b) { return len(a) > len(b); });
  std::vector<int> pos(n);
  for (int p = 0; p < n; ++p) pos[row[p]] = p;
  const size_t nnz = mesh.col.size();
  std::vector<int> rp(n + 1, 0), col(nnz), dg(n, -1), src(nnz);
  size_t k = 0;
  std::vector<std::pair<int, int>> ent;
  for (int p = 0; p < n; ++p) {
    const int r = row[p];
    ent.clear();
    for (int q = mesh.rowptr[r]; q < mesh.rowptr[r + 1]; ++q) ent.emplace_back(pos[mesh.col[q]], q);
    std::sort(ent.begin(), ent.end());
    for (const auto& e : ent) {
      col[k] = e.first;
      src[k] = e.second;
      if (e.first == p) dg[p] = static_cast<int>(k);
      ++k;
    }
    rp[p + 1] = static_cast<int>(k);
  }
  std::vector<int> ptr[kSellVariants], srow[kSellVariants], ssrc[kSellVariants];
  std::vector<int16_t> off[kSellVariants];
  bool ident[kSellVariants];
  for (int v = 0; v < kSellVariants; ++v) {
    std::vector<int> scol;
    sell_pattern(n, rp.data(), col.data(), kSellSigma[v], ptr[v], srow[v], scol, ssrc[v]);
    if (!sell_offsets16(n, ptr[v], srow[v], scol, off[v])) return;  // stays natural
    ident[v] = true;
    for (int p = 0; p < n && ident[v]; ++p) ident[v] = srow[v][p] == p;
  }
  ord_pos_ = upload(allocs_, pos, stream);
  ord_row_ = upload(allocs_, row, stream);
  ord_rowptr_ = upload(allocs_, rp, stream);
  ord_col_ = upload(allocs_, col, stream);
  ord_diag_ = upload(allocs_, dg, stream);
  ord_src_ = upload(allocs_, src, stream);
  a32o_ = alloc<float>(nnz);
  for (int v = 0; v < kSellVariants; ++v) {
    SellPat pt;
    pt.slots = ssrc[v].size();
    pt.ptr = upload(allocs_, ptr[v], stream);
    pt.row = ident[v] ? nullptr : upload(allocs_, srow[v], stream);
    pt.col16 = upload(allocs_, off[v], stream);
    pt.src = upload(allocs_, ssrc[v], stream);  // positions in the ordered CSR
    if (ncase > 1) pt.scase = upload(allocs_, slice_cases(ptr[v], srow[v]), stream);
    osell_[v] = pt;
  }
  sync();  // host staging vectors go out of scope
  order_ok_ = true;
}

// The FP32 values of the ordered CSR from the FP64 matrix, with the Jacobi
// scaling folded in when that is the preconditioner
void Context::fill_order(const double* csr_val, bool jacobi) {
  fill_ordered(dm.Nc, ord_rowptr_, ord_src_, ord_row_, dm.diag, csr_val, jacobi, a32o_, stream);
  order_jacobi_ = jacobi;
}

// The FP32 workspace of a mixed GMRES solve in the inner order (after
// fill_order): the shapes of the FP32 workspace, the ordered SELL copies (of
// the window in use, or all of them) filled from the ordered CSR, and the
// position maps.
KWork Context::ordered_work(bool all) {
  KWork w = kwork32_;
  for (int v = 0; v < kSellVariants; ++v) {
    w.sell[v] = osell_[v];
    w.sell_vals[v] = nullptr;
    if (all || (w.lanes == 0 && v == w.sell_v)) {
      if (!osell32_[v]) osell32_[v] = alloc<float>(osell_[v].slots);
      sell_fill<float, float>(osell_[v].src, osell_[v].slots, a32o_, osell32_[v], stream);
      w.sell_vals[v] = osell32_[v];
    }
  }
  w.ord_pos = ord_pos_;
  w.ord_row = ord_row_;
  w.pc_folded = order_jacobi_;
  return w;
}

// The autotuner on the ordered system; the shapes it picks become the FP32
// workspace's.
void Context::tune_ordered(const CsrView<float>& Ao, int m) {
  KWork w = ordered_work(true);
  autotune<float>(Ao, w, m, stream);
  kwork32_.lanes = w.lanes;
  kwork32_.nt = w.nt;
  kwork32_.rows = w.rows;
  kwork32_.sell_v = w.sell_v;
  kwork32_.sell2 = w.sell2;
  kwork32_.report = w.report;
  kwork32_.tuned = true;
}

// Values of the SELL copies from the CSR values: of the window in use, or of
// every window (before the autotuner times them).
void Context::prepare_sell(const double* csr_val, bool fp32, bool all) {
  ensure_sell();
  KWork& w = fp32 ? kwork32_ : kwork_;
  for (int v = 0; v < kSellVariants; ++v) {
    if (!all && v != w.sell_v) continue;
    const SellPat& pt = w.sell[v];
    if (fp32) {
      if (!sell32_[v]) sell32_[v] = alloc<float>(pt.slots);
      sell_fill<float, double>(pt.src, pt.slots, csr_val, sell32_[v], stream);
      w.sell_vals[v] = sell32_[v];
    } else {
      if (!sell64_[v]) sell64_[v] = alloc<double>(pt.slots);
      sell_fill<double, double>(pt.src, pt.slots, csr_val, sell64_[v], stream);
      w.sell_vals[v] = sell64_[v];
    }
  }
}

double Context::time_spmv(int which, int precision, int repeats) {
  const double* val = which == 0 ? df.Au : (which == 1 ? df.Ap : df.As);
  if (repeats < 1) repeats = 1;
  const int m = std::min(par.ksp[S355_KSP_PRESSURE].restart, kMaxRestart);
  // a scratch buffer larger than the 256 MB Infinity Cache
  const size_t flush_bytes = size_t(512) << 20;
  void* flush = nullptr;
  S355_HIP(hipMalloc(&flush, flush_bytes));
  double ms = 0.0;
  try {
    if (precision == 0) {
      CsrView<double> A{dm.Nc, dm.rowptr, dm.col, dm.diag, val, dm.nnz};
      prepare_sell(val, false, par.autotune && !kwork_.tuned);
      if (par.autotune && !kwork_.tuned) autotune<double>(A, kwork_, m, stream);
      double* inv = static_cast<double*>(kwork_.pcinv);
      jacobi_inverse<double>(A, inv, stream);
      double* y = static_cast<double*>(kwork_.V);
      ms = bench_spmv<double>(A, df.U, y, inv, kwork_, repeats, flush, flush_bytes, stream);
    } else {
      ensure_mixed();
      convert_to_float(val, a32_, dm.nnz, stream);
      CsrView<float> A{dm.Nc, dm.rowptr, dm.col, dm.diag, a32_, dm.nnz};
      float* inv = static_cast<float*>(kwork32_.pcinv);
      float* xf = static_cast<float*>(kwork32_.V);
      float* yf = xf + dm.Nc;
      s355_ksp_opts o = par.ksp[S355_KSP_PRESSURE];
      o.type = S355_KSP_GMRES;
      const bool tune = par.autotune && !kwork32_.tuned;
      if (use_order(o)) {
        // the product the pressure solve runs: the inner order's, its
        // Jacobi scaling in the matrix
        fill_order(val, o.pc == S355_PC_JACOBI);
        CsrView<float> Ao{dm.Nc, ord_rowptr_, ord_col_, ord_diag_, a32o_, dm.nnz};
        if (tune) tune_ordered(Ao, m);
        KWork w = ordered_work(false);
        convert_to_float(df.U, xf, dm.Nc, stream);
        jacobi_inverse<float>(Ao, inv, stream);
        ms = bench_spmv<float>(Ao, xf, yf, w.pc_folded ? nullptr : inv, w, repeats, flush, flush_bytes, stream);
      } else {
        prepare_sell(val, true, tune);
        if (tune) autotune<float>(A, kwork32_, m, stream);
        convert_to_float(df.U, xf, dm.Nc, stream);
        jacobi_inverse<float>(A, inv, stream);
        ms = bench_spmv<float>(A, xf, yf, inv, kwork32_, repeats, flush, flush_bytes, stream);
      }
    }
  } catch (...) {
    hipStreamSynchronize(stream);
    hipFree(flush);
    throw;
  }
  sync();
  hipFree(flush);
  return ms;
}

double Context::spmv_profile(int precision, int mode, int* count) {
  if (precision < 0 || precision > 1 || mode < -1 || mode > 1) throw std::invalid_argument("spmv_profile");
  KWork& w = precision ? kwork32_ : kwork_;
  std::unique_ptr<SpmvTimer>& t = timers_[precision];
  if (mode == 1) {
    t = std::make_unique<SpmvTimer>();
    w.timer = t.get();
  }
  double avg = 0.0;
  int n = 0;
  if (t) {
    sync();
    t->harvest();
    n = static_cast<int>(t->count);
    avg = t->count ? t->total_ms / t->count : 0.0;
  }
  if (mode == -1) w.timer = nullptr;
  if (count) *count = n;
  return avg;
}

void Context::set_kernel_shapes(int precision, int lanes, int nt, int rows, int sigma) {
  auto ok = [](int v, std::initializer_list<int> set) {
    for (int x : set)
      if (v == x) return true;
    return false;
  };
  if (precision < 0 || precision > 1 || !ok(lanes, {-2, -1, 0, 1, 2, 4, 8, 16}) || !ok(nt, {0, 1}) ||
      !ok(rows, {0, 1, 2, 4, 8, 16}) || !ok(sigma, {0, kSellSigma[0], kSellSigma[1], kSellSigma[2]}) ||
      (lanes == -2 && (precision != 1 || ncase != 1)))
    throw std::invalid_argument(
        "kernel shapes: lanes 0 (SELL-64), -1 (SELL-64, one slice per wave), -2 (SELL-64, two slices per wave: FP32, "
        "one case), 1, 2, 4, 8, 16; nt 0/1; rows 0, 1, 2, 4, 8, 16; sigma 0, 64, 512, 4096");
  KWork& w = precision ? kwork32_ : kwork_;
  w.lanes = lanes < 0 ? 0 : lanes;
  w.nt = nt;
  w.rows = rows;
  w.sell_v = kSellDefault;
  for (int v = 0; v < kSellVariants; ++v)
    if (sigma == kSellSigma[v]) w.sell_v = v;
  // FP32 SELL: two slices per wave above 2^20 rows, or as forced
  w.sell2 = lanes == -2 ? 1 : (lanes == -1 ? 0 : -1);
  w.tuned = true;  // the autotuner leaves explicit shapes alone
}

void Context::kernel_shapes(int lanes[2], int nt[2], int rows[2], int sigma[2], int tuned[2], int slices[2],
                            int order[2]) {
  order[0] = 0;
  s355_ksp_opts o = par.ksp[S355_KSP_PRESSURE];
  o.type = S355_KSP_GMRES;
  order[1] = use_order(o) ? 1 : 0;
  const KWork* w[2] = {&kwork_, &kwork32_};
  const size_t elem[2] = {sizeof(double), sizeof(float)};
  for (int k = 0; k < 2; ++k) {
    lanes[k] = w[k]->lanes;
    nt[k] = w[k]->nt;
    rows[k] = w[k]->rows ? w[k]->rows : pick_rows(dm.Nc, elem[k]);
    sigma[k] = kSellSigma[w[k]->sell_v];
    tuned[k] = w[k]->tuned ? 1 : 0;
    // the rule of the SELL launch: FP32 on one case, as tuned or above 2^20 rows
    const bool two = k == 1 && ncase == 1 && (w[k]->sell2 > 0 || (w[k]->sell2 < 0 && dm.Nc >= (1 << 20)));
    slices[k] = lanes[k] != 0 ? 0 : (two ? 2 : 1);
  }
}

double Context::time_copy(size_t bytes, int repeats) {
  if (bytes == 0 || bytes % 16) throw std::invalid_argument("copy size must be a positive multiple of 16 bytes");
  if (repeats < 1) repeats = 1;
  void *a = nullptr, *b = nullptr;
  S355_HIP(hipMalloc(&a, bytes));
  if (hipMalloc(&b, bytes) != hipSuccess) {
    hipFree(a);
    throw HipError("time_copy: out of device memory");
  }
  double ms = 0.0;
  try {
    fill_zero(a, bytes, stream);
    ms = bench_copy(b, a, bytes, repeats, stream);
  } catch (...) {
    hipStreamSynchronize(stream);
    hipFree(a);
    hipFree(b);
    throw;
  }
  S355_HIP(hipStreamSynchronize(stream));
  hipFree(a);
  hipFree(b);
  return ms;
}

hipEvent_t Context::take_event() {
  if (!pool_.empty()) {
    hipEvent_t e = pool_.back();
    pool_.pop_back();
    return e;
  }
  hipEvent_t e;
  S355_HIP(hipEventCreate(&e));
  all_events_.push_back(e);
  return e;
}

void Context::stage_push(int stage) {
  if (stage < 1 || stage > 5) throw std::invalid_argument("log stage must be 1..5");
  hipEvent_t a = take_event();
  S355_HIP(hipEventRecord(a, stream));
  open_.push_back({stage, a, nullptr});
}

void Context::stage_pop() {
  if (open_.empty()) throw std::logic_error("log stage pop without push");
  StageEvent ev = open_.back();
  open_.pop_back();
  ev.b = take_event();
  S355_HIP(hipEventRecord(ev.b, stream));
  done_.push_back(ev);
  if (done_.size() > 512) resolve_stages();
}

void Context::resolve_stages() {
  for (auto& ev : done_) {
    S355_HIP(hipEventSynchronize(ev.b));
    float ms = 0.f;
    S355_HIP(hipEventElapsedTime(&ms, ev.a, ev.b));
    stage_ms_[ev.stage] += ms;
    stage_ms_[0] += ms;
    pool_.push_back(ev.a);
    pool_.push_back(ev.b);
  }
  done_.clear();
}

void Context::stage_times(double ms[6], bool reset) {
  resolve_stages();
  for (int k = 0; k < 6; ++k) ms[k] = stage_ms_[k];
  if (reset)
    for (double& v : stage_ms_) v = 0.0;
}

}  // namespace s355
