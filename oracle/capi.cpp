// ORACLE — test infrastructure only (see oracle.hpp).
//
// C entry points of the oracle, mirroring include/simple355.h one for one
// with an orc_ prefix so that tests drive both implementations the same way.
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle.hpp"

namespace {
thread_local std::string g_err;

int fail(const std::exception& e) {
  g_err = e.what();
  return S355_E_ARG;
}

orc::Oracle* O(void* h) { return static_cast<orc::Oracle*>(h); }

size_t field_len(orc::Oracle* o, int id) {
  std::vector<double>* v = o->field(id);
  return v ? v->size() : 0;
}

void fill_info(const orc::SolveResult& r, s355_solve_info* info) {
  info->iters = r.iters;
  info->reason = r.reason;
  info->rnorm = r.rnorm;
  info->rnorm0 = r.rnorm0;
  info->refinements = 0;
  info->pad = 0;
}

template <class F>
int guard(F&& f) {
  try {
    return f();
  } catch (const std::exception& e) {
    return fail(e);
  }
}
}  // namespace

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int orc_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void* orc_create(const s355_mesh_desc* mesh, const s355_params* params) {
  try {
    return new orc::Oracle(*mesh, *params);
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

void orc_destroy(void* h) { delete O(h); }

int orc_sizes(void* h, s355_sizes* s) {
  orc::Oracle* o = O(h);
  s->n_vert = o->Nvrt;
  s->n_cell = o->Ncel;
  s->n_face = o->Nfac;
  s->n_bnd = o->Nbnd;
  s->n_interior = o->Nfac - o->Nbnd;
  s->nnz = o->csr.nnz();
  s->n_elem = static_cast<int>(o->perm.size());
  s->n_warped = o->n_warped;
  return S355_OK;
}

int orc_get_mesh_int(void* h, int which, int* out, size_t n) {
  return guard([&] {
    orc::Oracle* o = O(h);
    std::vector<int> v;
    const int K = orc::kMaxFace;
    switch (which) {
      case S355_M_ELEM_PERM: v = o->perm; break;
      case S355_M_FACE_C1: for (auto& f : o->face) v.push_back(f.c1); break;
      case S355_M_FACE_C2: for (auto& f : o->face) v.push_back(f.c2); break;
      case S355_M_FACE_BND: for (auto& f : o->face) v.push_back(f.bnd); break;
      case S355_M_FACE_VERT:
        for (auto& f : o->face) v.insert(v.end(), f.v, f.v + 4);
        break;
      case S355_M_CELL_NFACE: for (auto& c : o->cell) v.push_back(c.nface); break;
      case S355_M_CELL_FACE: for (auto& c : o->cell) v.insert(v.end(), c.face, c.face + K); break;
      case S355_M_CELL_CELL: for (auto& c : o->cell) v.insert(v.end(), c.cell, c.cell + K); break;
      case S355_M_BND_FACE: for (auto& b : o->bnd) v.push_back(b.face); break;
      case S355_M_BND_RID: for (auto& b : o->bnd) v.push_back(b.rid); break;
      case S355_M_CELL_RID: for (auto& c : o->cell) v.push_back(c.rid); break;
      case S355_M_CSR_ROWPTR: v = o->csr.rowptr; break;
      case S355_M_CSR_COL: v = o->csr.col; break;
      case S355_M_CSR_DIAG: v = o->csr.diag; break;
      case S355_M_CELL_VERT: for (auto& c : o->cell) v.insert(v.end(), c.v, c.v + 8); break;
      default: throw std::runtime_error("bad mesh int id");
    }
    if (n != v.size()) throw std::runtime_error("size mismatch for mesh int " + std::to_string(which));
    std::memcpy(out, v.data(), n * sizeof(int));
    return S355_OK;
  });
}

int orc_get_mesh_dbl(void* h, int which, double* out, size_t n) {
  return guard([&] {
    orc::Oracle* o = O(h);
    std::vector<double> v;
    switch (which) {
      case S355_G_CELL_VOL: for (auto& c : o->cell) v.push_back(c.vol); break;
      case S355_G_CELL_X: for (auto& c : o->cell) v.insert(v.end(), c.x, c.x + 3); break;
      case S355_G_FACE_X: for (auto& f : o->face) v.insert(v.end(), f.x, f.x + 3); break;
      case S355_G_FACE_N: for (auto& f : o->face) v.insert(v.end(), f.n, f.n + 3); break;
      case S355_G_FACE_AREA: for (auto& f : o->face) v.push_back(f.area); break;
      case S355_G_FACE_LAMBDA: for (auto& f : o->face) v.push_back(f.lambda); break;
      case S355_G_FACE_RLENCOS: for (auto& f : o->face) v.push_back(f.rlencos); break;
      case S355_G_FACE_XPAC: for (auto& f : o->face) v.insert(v.end(), f.xpac, f.xpac + 3); break;
      case S355_G_FACE_XNAC: for (auto& f : o->face) v.insert(v.end(), f.xnac, f.xnac + 3); break;
      case S355_G_BND_DIST: for (auto& b : o->bnd) v.push_back(b.distance); break;
      case S355_G_VERT: v = o->vert; break;
      default: throw std::runtime_error("bad mesh dbl id");
    }
    if (n != v.size()) throw std::runtime_error("size mismatch for mesh dbl " + std::to_string(which));
    std::memcpy(out, v.data(), n * sizeof(double));
    return S355_OK;
  });
}

int orc_set_field(void* h, int id, const double* in, size_t n) {
  return guard([&] {
    std::vector<double>* v = O(h)->field(id);
    if (!v || v->size() != n) throw std::runtime_error("bad field id or size");
    std::memcpy(v->data(), in, n * sizeof(double));
    return S355_OK;
  });
}

int orc_get_field(void* h, int id, double* out, size_t n) {
  return guard([&] {
    std::vector<double>* v = O(h)->field(id);
    if (!v || v->size() != n) throw std::runtime_error("bad field id or size");
    std::memcpy(out, v->data(), n * sizeof(double));
    return S355_OK;
  });
}

size_t orc_field_len(void* h, int id) { return field_len(O(h), id); }

int orc_init_flow_field(void* h) { return guard([&] { O(h)->InitFlowField(); return S355_OK; }); }
int orc_bc_velocity(void* h) { return guard([&] { O(h)->SetBCVelocity(); return S355_OK; }); }
int orc_bc_pressure(void* h) { return guard([&] { O(h)->SetBCPressure(); return S355_OK; }); }
int orc_bc_deltap(void* h) { return guard([&] { O(h)->SetBCDeltaP(); return S355_OK; }); }

int orc_gradient(void* h, int field, int bfield, int out) {
  return guard([&] {
    orc::Oracle* o = O(h);
    std::vector<double>*f = o->field(field), *b = o->field(bfield), *g = o->field(out);
    if (!f || !b || !g) throw std::runtime_error("bad field id");
    if (f->size() != static_cast<size_t>(o->Ncel) || b->size() != static_cast<size_t>(o->Nbnd) ||
        g->size() != 3 * static_cast<size_t>(o->Ncel))
      throw std::runtime_error("gradient: field shapes");
    o->Gradient(*f, *b, *g);
    return S355_OK;
  });
}

int orc_build_velocity_matrix(void* h) { return guard([&] { O(h)->BuildVelocityMatrix(); return S355_OK; }); }

int orc_solve_velocity(void* h, double rtol, int maxit, s355_solve_info info[3]) {
  return guard([&] {
    int rc = O(h)->SolveVelocity(rtol, maxit);
    for (int k = 0; k < 3; ++k) fill_info(O(h)->last[k], &info[k]);
    return rc;
  });
}

int orc_cal_ru_face(void* h) { return guard([&] { O(h)->CalRUFace2(); return S355_OK; }); }
int orc_build_pressure_matrix(void* h) { return guard([&] { O(h)->BuildPressureMatrix(); return S355_OK; }); }

int orc_solve_pressure(void* h, double rtol, int maxit, s355_solve_info* info) {
  return guard([&] {
    int rc = O(h)->SolvePressure(rtol, maxit);
    fill_info(O(h)->last[3], info);
    return rc;
  });
}

int orc_correct_fields(void* h, double* res_p) {
  return guard([&] {
    double r = 0.0;
    O(h)->CorrectFields(&r);
    *res_p = r;
    return S355_OK;
  });
}

int orc_correct_ru_face(void* h) { return guard([&] { O(h)->CorrectRUFace2(); return S355_OK; }); }

int orc_calculate_velocity(void* h, double res[3], s355_solve_info info[3]) {
  return guard([&] {
    orc::Oracle* o = O(h);
    for (int k = 0; k < 3; ++k) o->localRes[k] = 0.0;
    int rc = o->CalculateVelocity();
    for (int k = 0; k < 3; ++k) {
      res[k] = o->localRes[k];
      fill_info(o->last[k], &info[k]);
    }
    return rc;
  });
}

int orc_calculate_pressure(void* h, double* res_p, s355_solve_info* info) {
  return guard([&] {
    orc::Oracle* o = O(h);
    o->localRes[3] = 0.0;
    int rc = o->CalculatePressure();
    *res_p = o->localRes[3];
    fill_info(o->last[3], info);
    return rc;
  });
}

int orc_update_tur_kepsilon(void* h, s355_solve_info info[2]) {
  return guard([&] {
    orc::Oracle* o = O(h);
    int rc = o->UpdateTurKEpsilon();
    fill_info(o->last[4], &info[0]);
    fill_info(o->last[5], &info[1]);
    return rc;
  });
}

// z = M^-1 r for a caller's CSR matrix (diag: the diagonal's position in
// each row) with the Jacobi or ILU(0) preconditioner of the Krylov solves
int orc_pc_apply(int n, const int* rowptr, const int* col, const int* diag, const double* val, int pc,
                 const double* r, double* z) {
  return guard([&] {
    if (n < 0 || !rowptr || !col || !diag || !val || !r || !z) throw std::invalid_argument("pc_apply: bad arguments");
    orc::Csr A;
    A.rowptr.assign(rowptr, rowptr + n + 1);
    A.col.assign(col, col + rowptr[n]);
    A.diag.assign(diag, diag + n);
    std::vector<double> v(val, val + rowptr[n]), rr(r, r + n), zz;
    orc::pc_apply(A, v, pc, rr, zz);
    std::memcpy(z, zz.data(), sizeof(double) * static_cast<size_t>(n));
    return S355_OK;
  });
}

// the pieces of UpdateTurKEpsilon, one entry each as in the device ABI
int orc_kepsilon_sources(void* h) { return guard([&] { O(h)->KEpsilonSources(); return S355_OK; }); }
int orc_bc_kepsilon(void* h) { return guard([&] { O(h)->KEpsilonBC(); return S355_OK; }); }
int orc_kepsilon_time_terms(void* h) { return guard([&] { O(h)->KEpsilonTimeTerms(); return S355_OK; }); }
int orc_build_scalar_matrix(void* h, int isca) {
  return guard([&] { O(h)->BuildKEpsilonMatrix(isca); return S355_OK; });
}
int orc_eps_wall_rows(void* h) { return guard([&] { O(h)->EpsWallRows(); return S355_OK; }); }
int orc_kepsilon_post_k(void* h) { return guard([&] { O(h)->KEpsilonPostK(); return S355_OK; }); }
int orc_kepsilon_post_eps(void* h) { return guard([&] { O(h)->KEpsilonPostEps(); return S355_OK; }); }
int orc_solve_scalar(void* h, int isca, double rtol, int maxit, s355_solve_info* info) {
  return guard([&] {
    if (isca < 1 || isca > 3) throw std::invalid_argument("solve_scalar: isca 1, 2 or 3");
    int rc = O(h)->SolveScalar(isca, rtol, maxit);
    fill_info(O(h)->last[isca == 1 ? 6 : isca + 2], info);
    return rc;
  });
}

int orc_bc_temperature(void* h) { return guard([&] { O(h)->SetBCTemperature(); return S355_OK; }); }

int orc_update_energy(void* h, s355_solve_info* info) {
  return guard([&] {
    int rc = O(h)->UpdateEnergy();
    if (info) fill_info(O(h)->last[6], info);
    return rc;
  });
}

int orc_simple_iterate(void* h, int n, double* history, int* done) {
  return guard([&] { return O(h)->SimpleIterate(n, history, done); });
}

int orc_save_transient_old_data(void* h) { return guard([&] { O(h)->SaveTransientOldData(); return S355_OK; }); }

int orc_time_state(void* h, double* cur_time, int* time_steps) {
  return guard([&] {
    *cur_time = O(h)->cur_time;
    *time_steps = O(h)->time_step;
    return S355_OK;
  });
}

}  // extern "C"
