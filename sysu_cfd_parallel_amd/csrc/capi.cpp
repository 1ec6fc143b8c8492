// extern "C" boundary of the library (include/simple355.h).  Each entry
// catches every exception and turns it into a return code plus a message
// for s355_last_error(); nothing blocks or aborts the process (the
// reference's ErrorHandler waits on getchar() then aborts, BasicType.h:23-55).
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/simple355.h"
#include "device/common.hpp"
#include "device/context.hpp"
#include "host/export.hpp"
#include "host/navier.hpp"
#include "host/output.hpp"
#include "mesh/cases.hpp"
#include "mesh/gmsh.hpp"
#include "mesh/ingest.hpp"

struct s355_ctx {
  std::unique_ptr<s355::Context> dev;
  std::unique_ptr<s355host::NavierStokesSolver> ns;
  std::unique_ptr<s355host::FieldExport> out;  // created on first use; released before dev
};

struct s355_case {
  s355::ElementStream es;
};

struct s355_mesh {
  s355::Mesh m;
};

namespace {

thread_local std::string g_err;

template <class F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const s355host::ConvergeError& e) {
    g_err = e.varname + " not converged in " + std::to_string(e.iter) + " iterations, residual " +
            std::to_string(e.residual) + ", reason " + std::to_string(e.reason);
    return S355_NOT_CONVERGED;
  } catch (const s355host::DeviceError& e) {
    g_err = e.msg;
    return e.code;
  } catch (const s355::HipError& e) {
    g_err = e.what();
    return S355_E_HIP;
  } catch (const s355host::IoError& e) {
    g_err = e.what();
    return S355_E_IO;
  } catch (const std::invalid_argument& e) {
    g_err = e.what();
    return S355_E_ARG;
  } catch (const std::bad_alloc&) {
    g_err = "host allocation failed";
    return S355_E_NOMEM;
  } catch (const std::logic_error& e) {
    g_err = e.what();
    return S355_E_STATE;
  } catch (const std::runtime_error& e) {
    g_err = e.what();
    return S355_E_MESH;
  } catch (...) {
    g_err = "unknown error";
    return S355_E_STATE;
  }
}

s355::Context& dev(s355_ctx* c) {
  if (!c || !c->dev) throw std::invalid_argument("null context");
  return *c->dev;
}

void sizes_of(const s355::Mesh& m, s355_sizes* s) {
  s->n_vert = m.n_vert;
  s->n_cell = m.n_cell;
  s->n_face = m.n_face;
  s->n_bnd = m.n_bnd;
  s->n_interior = m.n_face - m.n_bnd;
  s->nnz = m.nnz();
  s->n_elem = m.n_elem;
  s->n_warped = m.n_warped;
}

int mesh_int(const s355::Mesh& m, int which, int* out, size_t n) {
  const s355::hvec<int>* v = nullptr;
  s355::hvec<int> tmp;
  switch (which) {
    case S355_M_ELEM_PERM: v = &m.perm; break;
    case S355_M_FACE_C1: v = &m.f_c1; break;
    case S355_M_FACE_C2: v = &m.f_c2; break;
    case S355_M_FACE_BND: v = &m.f_bnd; break;
    case S355_M_FACE_VERT: v = &m.f_v; break;
    case S355_M_CELL_NFACE: v = &m.c_nface; break;
    case S355_M_CELL_FACE: v = &m.c_face; break;
    case S355_M_CELL_CELL: v = &m.c_cell; break;
    case S355_M_BND_FACE:
      tmp.resize(m.n_bnd);
      for (int b = 0; b < m.n_bnd; ++b) tmp[b] = b;  // boundary faces come first (geometry.cpp:250-265)
      v = &tmp;
      break;
    case S355_M_BND_RID: v = &m.b_rid; break;
    case S355_M_CELL_RID: v = &m.c_rid; break;
    case S355_M_CSR_ROWPTR: v = &m.rowptr; break;
    case S355_M_CSR_COL: v = &m.col; break;
    case S355_M_CSR_DIAG: v = &m.diag; break;
    case S355_M_CELL_VERT: v = &m.c_vert; break;
    default: throw std::invalid_argument("unknown mesh int id");
  }
  if (!out || n != v->size()) throw std::invalid_argument("mesh int " + std::to_string(which) + ": expected length " +
                                                          std::to_string(v->size()));
  std::memcpy(out, v->data(), n * sizeof(int));
  return S355_OK;
}

int mesh_dbl(const s355::Mesh& m, int which, double* out, size_t n) {
  const s355::hvec<double>* v = nullptr;
  switch (which) {
    case S355_G_CELL_VOL: v = &m.c_vol; break;
    case S355_G_CELL_X: v = &m.c_x; break;
    case S355_G_FACE_X: v = &m.f_x; break;
    case S355_G_FACE_N: v = &m.f_n; break;
    case S355_G_FACE_AREA: v = &m.f_area; break;
    case S355_G_FACE_LAMBDA: v = &m.f_lambda; break;
    case S355_G_FACE_RLENCOS: v = &m.f_rlencos; break;
    case S355_G_FACE_XPAC: v = &m.f_xpac; break;
    case S355_G_FACE_XNAC: v = &m.f_xnac; break;
    case S355_G_BND_DIST: v = &m.b_dist; break;
    case S355_G_VERT: v = &m.vert; break;
    default: throw std::invalid_argument("unknown mesh dbl id");
  }
  if (!out || n != v->size()) throw std::invalid_argument("mesh dbl " + std::to_string(which) + ": expected length " +
                                                          std::to_string(v->size()));
  std::memcpy(out, v->data(), n * sizeof(double));
  return S355_OK;
}

void check_params(const s355_params* p) {
  if (!p) throw std::invalid_argument("null params");
  for (int k = 0; k < 3; ++k) {
    const s355_ksp_opts& o = p->ksp[k];
    if (o.type != S355_KSP_GMRES && o.type != S355_KSP_CG) throw std::invalid_argument("unknown KSP type");
    if (o.pc < S355_PC_NONE || o.pc > S355_PC_ILU0) throw std::invalid_argument("unknown PC type");
    if (o.restart < 1 || o.restart > s355::kMaxRestart) throw std::invalid_argument("GMRES restart must be 1..64");
    if (o.maxit < 1 || !(o.rtol >= 0.0) || !(o.dtol > 0.0)) throw std::invalid_argument("bad KSP tolerances");
  }
  if (!(p->urf[0] > 0.0) || !(p->urf[5] > 0.0)) throw std::invalid_argument("relaxation factors must be positive");
  if (p->solve_energy && !(p->cp > 0.0 && p->prl > 0.0 && p->prte > 0.0))
    throw std::invalid_argument("energy: cp, prl and prte must be positive");
  if (p->limiter < 0 || p->limiter > 3) throw std::invalid_argument("limiter: 0 none, 1 Barth, 2 MLP, 3 WENO");
  if (p->steady != 0 && p->steady != 1) throw std::invalid_argument("steady: 0 or 1");
  if (!p->steady) {
    if (p->time_scheme != 1 && p->time_scheme != 2) throw std::invalid_argument("time_scheme: 1 Euler, 2 BDF2");
    if (!(p->dt > 0.0)) throw std::invalid_argument("dt must be positive");
    if (p->max_outer_step < 2) throw std::invalid_argument("max_outer_step must be at least 2");
  }
}

// the message of a solve that ended without converging (its rc is not an
// exception, so guarded() does not set one)
void note_not_converged(const char* var, size_t len, const s355_solve_info& i) {
  g_err = std::string(var, strnlen(var, len)) + " did not converge: reason " + std::to_string(i.reason) + " after " +
          std::to_string(i.iters) + " iterations, residual " + std::to_string(i.rnorm);
}

}  // namespace

extern "C" {

const char* s355_last_error(void) { return g_err.c_str(); }

void s355_default_params(s355_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->turb_model = 0;
  p->density_model = 0;
  p->steady = 1;
  p->cell_pressure_ref = 0;
  const double urf[8] = {0.6, 0.6, 0.6, 0.5, 0.8, 0.8, 0.8, 0.8};
  for (int k = 0; k < 8; ++k) p->urf[k] = urf[k];
  p->residual_steady = 1.e-6;
  // in-code solve settings on PETSc defaults; the pressure solve asks for CG
  // because the matrix is flagged symmetric (MPIStructure.cpp:462-465)
  p->ksp[S355_KSP_VELOCITY] = s355_ksp_opts{S355_KSP_GMRES, 30, S355_PC_ILU0, 1000, 1.e-8, 1.e-50, 1.e5};
  p->ksp[S355_KSP_PRESSURE] = s355_ksp_opts{S355_KSP_CG, 30, S355_PC_ILU0, 1000, 1.e-6, 1.e-50, 1.e5};
  p->ksp[S355_KSP_SCALAR] = s355_ksp_opts{S355_KSP_GMRES, 30, S355_PC_ILU0, 500, 1.e-8, 1.e-50, 1.e5};
  p->precision = S355_FP64;
  p->ir_max = 30;
  p->ir_inner_rtol = 1.e-3;
  p->autotune = 0;
  p->graphs = 0;
  p->solve_energy = 0;  // navier.cpp:481
  p->cp = 1006.;        // navier.cpp:489-492
  p->prl = 0.72;
  p->prte = 0.9;
  p->limiter = 0;  // navier.cpp:509
  p->time_scheme = 1;  // navier.cpp:512
  p->dt = 1.0;  // navier.cpp:511
  p->total_time = 0.;
  p->max_outer_step = 50;  // navier.cpp:476
}

int s355_case_build(int kind, const int* ip, const double* dp, s355_case** out) {
  return guarded([&] {
    if (!ip || !dp || !out) throw std::invalid_argument("null argument");
    auto c = std::make_unique<s355_case>();
    if (kind == 0) {
      s355::CavitySpec s;
      s.n = ip[0];
      s.tets = ip[1] != 0;
      s.seed = static_cast<uint64_t>(ip[2]);
      s.re = dp[0];
      s.lid_u = dp[1];
      s.jitter = dp[2];
      c->es = s355::make_cavity(s);
    } else if (kind == 1) {
      s355::ChannelSpec s;
      s.nx = ip[0];
      s.ny = ip[1];
      s.nz = ip[2];
      s.turbulent = ip[3] != 0;
      s.seed = static_cast<uint64_t>(ip[4]);
      s.lx = dp[0];
      s.ly = dp[1];
      s.lz = dp[2];
      s.jitter = dp[3];
      s.split_frac = dp[4];
      s.re = dp[5];
      s.u_in = dp[6];
      s.intensity = dp[7];
      c->es = s355::make_channel(s);
    } else if (kind == 2) {
      s355::StepCaseSpec s;
      s.n_per_h = ip[0];
      s.tets = ip[1] != 0;
      s.turbulent = ip[2] != 0;
      s.seed = static_cast<uint64_t>(ip[3]);
      s.upstream = dp[0];
      s.downstream = dp[1];
      s.span = dp[2];
      s.jitter = dp[3];
      s.re = dp[4];
      s.u_in = dp[5];
      s.intensity = dp[6];
      c->es = s355::make_step_case(s);
    } else {
      throw std::invalid_argument("unknown case kind");
    }
    *out = c.release();
    return S355_OK;
  });
}

int s355_case_sizes(const s355_case* c, int* sizes4) {
  return guarded([&] {
    if (!c || !sizes4) throw std::invalid_argument("null argument");
    sizes4[0] = c->es.n_vert();
    sizes4[1] = c->es.n_elem();
    sizes4[2] = static_cast<int>(c->es.vidx.size());
    sizes4[3] = static_cast<int>(c->es.regions.size());
    return S355_OK;
  });
}

int s355_case_copy(const s355_case* c, double* xyz, int* type, int* tag, int* vptr, int* vidx,
                   s355_region* regions) {
  return guarded([&] {
    if (!c) throw std::invalid_argument("null case");
    const auto& e = c->es;
    if (xyz) std::memcpy(xyz, e.xyz.data(), e.xyz.size() * sizeof(double));
    if (type) std::memcpy(type, e.type.data(), e.type.size() * sizeof(int));
    if (tag) std::memcpy(tag, e.tag.data(), e.tag.size() * sizeof(int));
    if (vptr) std::memcpy(vptr, e.vptr.data(), e.vptr.size() * sizeof(int));
    if (vidx) std::memcpy(vidx, e.vidx.data(), e.vidx.size() * sizeof(int));
    if (regions)
      for (size_t r = 0; r < e.regions.size(); ++r) {
        const auto& g = e.regions[r];
        s355_region o{};
        o.rid = g.rid;
        o.type1 = g.type1;
        o.type2 = g.type2;
        o.fixed_value = g.fixed_value;
        for (int k = 0; k < 10; ++k) o.init[k] = g.init[k];
        regions[r] = o;
      }
    return S355_OK;
  });
}

void s355_case_free(s355_case* c) { delete c; }

int s355_case_read(const char* path, int binary, s355_case** out) {
  return guarded([&] {
    if (!path || !out) throw std::invalid_argument("null argument");
    auto c = std::make_unique<s355_case>();
    c->es = binary ? s355::read_gmsh_binary(path) : s355::read_gmsh_ascii(path);
    *out = c.release();
    return S355_OK;
  });
}

int s355_case_write(const s355_mesh_desc* mesh, const char* path, int binary) {
  return guarded([&] {
    if (!mesh || !path) throw std::invalid_argument("null argument");
    s355::ElementStream es;
    es.xyz.assign(mesh->xyz, mesh->xyz + 3 * static_cast<size_t>(mesh->n_vert));
    for (int e = 0; e < mesh->n_elem; ++e) {
      const int t = mesh->elem_type[e];
      if (s355::gmsh_nodes(t) != mesh->elem_vptr[e + 1] - mesh->elem_vptr[e])
        throw std::invalid_argument("element " + std::to_string(e) + ": vertex count does not match its type");
      es.add(t, mesh->elem_tag[e], mesh->elem_vidx + mesh->elem_vptr[e]);
    }
    if (binary)
      s355::write_gmsh_binary(es, path);
    else
      s355::write_gmsh_ascii(es, path);
    return S355_OK;
  });
}

int s355_mesh_build(const s355_mesh_desc* desc, s355_mesh** out) {
  return guarded([&] {
    if (!desc || !out) throw std::invalid_argument("null argument");
    auto m = std::make_unique<s355_mesh>();
    m->m = s355::ingest_mesh(*desc);
    *out = m.release();
    return S355_OK;
  });
}
int s355_mesh_build_batch(int n_case, const s355_mesh_desc* descs, s355_mesh** out) {
  return guarded([&] {
    if (!descs || !out || n_case < 1) throw std::invalid_argument("null argument or no cases");
    std::vector<s355::Mesh> parts;
    parts.reserve(n_case);
    for (int k = 0; k < n_case; ++k) parts.push_back(s355::ingest_mesh(descs[k]));
    auto m = std::make_unique<s355_mesh>();
    s355::BatchLayout lay;
    m->m = s355::concat_meshes(parts, lay);
    *out = m.release();
    return S355_OK;
  });
}
int s355_mesh_sizes(const s355_mesh* m, s355_sizes* out) {
  return guarded([&] {
    if (!m || !out) throw std::invalid_argument("null argument");
    sizes_of(m->m, out);
    return S355_OK;
  });
}
int s355_mesh_get_int(const s355_mesh* m, int which, int* out, size_t n) {
  return guarded([&] {
    if (!m) throw std::invalid_argument("null mesh");
    return mesh_int(m->m, which, out, n);
  });
}
int s355_mesh_get_dbl(const s355_mesh* m, int which, double* out, size_t n) {
  return guarded([&] {
    if (!m) throw std::invalid_argument("null mesh");
    return mesh_dbl(m->m, which, out, n);
  });
}
void s355_mesh_free(s355_mesh* m) { delete m; }

namespace {
s355host::OutFormat out_format(int format) {
  if (format < 0 || format >= S355_OUT_COUNT) throw std::invalid_argument("unknown output format " + std::to_string(format));
  return static_cast<s355host::OutFormat>(format);
}

// OutputMoniter's records (navierIO.cpp:405-468), summed on the host in the
// reference's order: boundary records, then cells, ascending
std::vector<s355host::MonitorRecord> monitor_records(s355::Context& d) {
  std::vector<s355_region> regs(d.mesh.regions.begin(), d.mesh.regions.end());
  std::sort(regs.begin(), regs.end(), [](const s355_region& a, const s355_region& b) { return a.rid < b.rid; });
  std::vector<s355host::MonitorRecord> rec;
  std::map<int, size_t> slot;
  std::map<int, const s355_region*> by_rid;
  for (const s355_region& r : regs) {
    by_rid[r.rid] = &r;
    if ((r.type1 != 1 && r.type1 != 5) || r.rid == 0) continue;
    slot[r.rid] = rec.size();
    s355host::MonitorRecord m;
    m.rid = r.rid;
    m.type1 = r.type1;
    rec.push_back(m);
  }
  const size_t nb = d.dm.Nb, nc = d.dm.Nc;
  std::vector<double> shear(3 * nb), w(nc);
  if (nb) S355_HIP(hipMemcpyAsync(shear.data(), d.df.shear, 3 * nb * sizeof(double), hipMemcpyDeviceToHost, d.stream));
  S355_HIP(hipMemcpyAsync(w.data(), d.df.W, nc * sizeof(double), hipMemcpyDeviceToHost, d.stream));
  d.sync();
  for (size_t i = 0; i < nb; ++i) {
    auto it = slot.find(d.mesh.b_rid[i]);
    if (it == slot.end() || rec[it->second].type1 != 1) continue;
    for (int k = 0; k < 3; ++k) rec[it->second].force[k] += shear[3 * i + k];
  }
  for (size_t i = 0; i < nc; ++i) {
    auto it = slot.find(d.mesh.c_rid[i]);
    if (it == slot.end() || rec[it->second].type1 != 5 || by_rid[d.mesh.c_rid[i]]->type2 != 0) continue;
    rec[it->second].scalar += w[i] * w[i] * d.mesh.c_vol[i];
  }
  return rec;
}

s355host::FieldExport& exporter(s355_ctx* c) {
  s355::Context& d = dev(c);
  if (!c->out) c->out = std::make_unique<s355host::FieldExport>(d);
  return *c->out;
}
}  // namespace

int s355_mesh_write_result(const s355_mesh* m, int format, const char* path, const s355_result* r) {
  return guarded([&] {
    if (!m || !path) throw std::invalid_argument("null argument");
    const s355host::OutFormat f = out_format(format);
    const bool grid = f == s355host::OutFormat::GridBinary || f == s355host::OutFormat::GridAscii;
    if (!grid && !r) throw std::invalid_argument("null result");
    s355host::ResultView v;
    v.n_vert = m->m.n_vert;
    v.n_cell = m->m.n_cell;
    v.vert = m->m.vert.data();
    v.cell_vert = m->m.c_vert.data();
    if (r) {
      v.R = r->R;
      v.U = r->U;
      v.V = r->V;
      v.W = r->W;
      v.P = r->P;
      v.T = r->T;
      v.vislam = r->vislam;
      v.vistur = r->vistur;
      v.te = r->te;
      v.ed = r->ed;
      v.turbulent = r->turbulent != 0;
    }
    s355host::write_result(v, f, path);
    return S355_OK;
  });
}

int s355_backup_read(const char* path, int n_cell, int turbulent, double* out) {
  return guarded([&] {
    if (!path || !out || n_cell < 1) throw std::invalid_argument("backup_read arguments");
    const size_t n = n_cell;
    s355host::BackupArrays a{out, out + n, out + 2 * n, out + 3 * n, out + 4 * n, out + 5 * n, out + 6 * n,
                             turbulent ? out + 7 * n : nullptr, turbulent ? out + 8 * n : nullptr,
                             turbulent ? out + 9 * n : nullptr};
    s355host::read_backup(path, n_cell, turbulent != 0, a);
    return S355_OK;
  });
}

int s355_output_snapshot(s355_ctx* ctx) {
  return guarded([&] {
    exporter(ctx).snapshot();
    return S355_OK;
  });
}

int s355_output_write(s355_ctx* ctx, int format, const char* path, int async) {
  return guarded([&] {
    if (!path) throw std::invalid_argument("null path");
    exporter(ctx).write({{out_format(format), path}}, async != 0);
    return S355_OK;
  });
}

int s355_output_wait(s355_ctx* ctx) {
  return guarded([&] {
    dev(ctx);
    if (ctx->out) ctx->out->wait();
    return S355_OK;
  });
}

int s355_read_backup(s355_ctx* ctx, const char* path) {
  return guarded([&] {
    s355::Context& d = dev(ctx);
    if (!path) throw std::invalid_argument("null path");
    if (ctx->out) ctx->out->wait();
    const bool turb = d.par.turb_model == 1;
    const size_t n = d.dm.Nc;
    std::vector<double> h((turb ? 10 : 7) * n);
    s355host::BackupArrays a{h.data(), h.data() + n, h.data() + 2 * n, h.data() + 3 * n, h.data() + 4 * n,
                             h.data() + 5 * n, h.data() + 6 * n, turb ? h.data() + 7 * n : nullptr,
                             turb ? h.data() + 8 * n : nullptr, turb ? h.data() + 9 * n : nullptr};
    s355host::read_backup(path, static_cast<int>(n), turb, a);
    const int ids[10] = {S355_F_R, S355_F_U, S355_F_V, S355_F_W, S355_F_P, S355_F_T, S355_F_VISLAM, S355_F_VISTUR,
                         S355_F_TE, S355_F_ED};
    for (int k = 0; k < (turb ? 10 : 7); ++k)
      S355_HIP(hipMemcpyAsync(d.field(ids[k]).p, h.data() + k * n, n * sizeof(double), hipMemcpyHostToDevice,
                              d.stream));
    d.sync();
    return S355_OK;
  });
}

int s355_monitor_records(s355_ctx* ctx, double* out, int max_records, int* n_records) {
  return guarded([&] {
    if (!out || !n_records) throw std::invalid_argument("null output");
    const auto rec = monitor_records(dev(ctx));
    if (static_cast<int>(rec.size()) > max_records) throw std::invalid_argument("monitor: more regions than max_records");
    for (size_t k = 0; k < rec.size(); ++k) {
      double* o = out + 6 * k;
      o[0] = rec[k].rid;
      o[1] = rec[k].type1;
      o[2] = rec[k].force[0];
      o[3] = rec[k].force[1];
      o[4] = rec[k].force[2];
      o[5] = rec[k].scalar;
    }
    *n_records = static_cast<int>(rec.size());
    return S355_OK;
  });
}

int s355_monitor_write(s355_ctx* ctx, const char* path, double cur_time, int header) {
  return guarded([&] {
    if (!path) throw std::invalid_argument("null path");
    const auto rec = monitor_records(dev(ctx));
    const std::string text = header ? s355host::monitor_header(rec) : s355host::monitor_line(cur_time, rec);
    FILE* f = std::fopen(path, header ? "wb" : "ab");
    if (!f) throw s355host::IoError(std::string("cannot open ") + path);
    const bool ok = std::fwrite(text.data(), 1, text.size(), f) == text.size();
    if (std::fclose(f) != 0 || !ok) throw s355host::IoError(std::string("error writing ") + path);
    return S355_OK;
  });
}

int s355_output_schedule(s355_ctx* ctx, int noutput, int binary, const char* dir) {
  return guarded([&] {
    dev(ctx);
    if (noutput < 0) throw std::invalid_argument("noutput < 0");
    if (noutput > 0 && !dir) throw std::invalid_argument("null output directory");
    ctx->ns->SetOutput(noutput, binary != 0, noutput > 0 ? dir : "");
    return S355_OK;
  });
}

int s355_ctx_create(const s355_mesh_desc* mesh, const s355_params* params, s355_ctx** out) {
  return guarded([&] {
    if (!mesh || !out) throw std::invalid_argument("null argument");
    *out = nullptr;
    check_params(params);
    auto c = std::make_unique<s355_ctx>();
    c->dev = std::make_unique<s355::Context>(mesh, 1, *params);
    c->ns = std::make_unique<s355host::NavierStokesSolver>(c.get(), *params);
    *out = c.release();
    return S355_OK;
  });
}

int s355_batch_create(int n_case, const s355_mesh_desc* meshes, const s355_params* params, s355_ctx** out) {
  return guarded([&] {
    if (!meshes || !out || n_case < 1) throw std::invalid_argument("null argument or no cases");
    *out = nullptr;
    check_params(params);
    auto c = std::make_unique<s355_ctx>();
    c->dev = std::make_unique<s355::Context>(meshes, n_case, *params);
    c->ns = std::make_unique<s355host::NavierStokesSolver>(c.get(), *params);
    *out = c.release();
    return S355_OK;
  });
}

int s355_batch_size(const s355_ctx* ctx, int* n_case) {
  return guarded([&] {
    if (!ctx || !n_case) throw std::invalid_argument("null argument");
    *n_case = ctx->dev->ncase;
    return S355_OK;
  });
}

int s355_batch_layout(const s355_ctx* ctx, int which, int* out, size_t n) {
  return guarded([&] {
    if (!ctx || !out) throw std::invalid_argument("null argument");
    const s355::BatchLayout& l = ctx->dev->lay;
    const std::vector<int>* v = nullptr;
    switch (which) {
      case S355_BATCH_CELL: v = &l.cell; break;
      case S355_BATCH_BND: v = &l.bnd; break;
      case S355_BATCH_IFACE: v = &l.iface; break;
      case S355_BATCH_NNZ: v = &l.nnz; break;
      case S355_BATCH_VERT: v = &l.vert; break;
      case S355_BATCH_ELEM: v = &l.elem; break;
      default: throw std::invalid_argument("unknown batch layout array");
    }
    if (n != v->size()) throw std::invalid_argument("batch layout: length must be n_case + 1");
    std::copy(v->begin(), v->end(), out);
    return S355_OK;
  });
}

int s355_batch_residuals(s355_ctx* ctx, double* res, int reset) {
  return guarded([&] {
    if (!res) throw std::invalid_argument("null res");
    s355::Context& d = dev(ctx);
    d.read_case_residuals(res);
    d.check_flags();
    if (reset) d.zero_residuals();
    return S355_OK;
  });
}

int s355_batch_solve_info(s355_ctx* ctx, int kind, s355_solve_info* info, int n_case) {
  return guarded([&] {
    s355::Context& d = dev(ctx);
    if (!info || n_case != d.ncase) throw std::invalid_argument("batch solve info: n_case entries");
    if (kind < 0 || kind > 6) throw std::invalid_argument("batch solve info: kind 0..6 (U, V, W, p', k, eps, T)");
    const auto& v = d.case_info[kind];
    for (int k = 0; k < n_case; ++k) {
      s355_solve_info x{};
      if (k < static_cast<int>(v.size())) {
        x.iters = v[k].iters;
        x.reason = v[k].reason;
        x.rnorm = v[k].rnorm;
        x.rnorm0 = v[k].rnorm0;
        x.refinements = v[k].refinements;
      }
      info[k] = x;
    }
    return S355_OK;
  });
}

int s355_ctx_destroy(s355_ctx* ctx) {
  return guarded([&] {
    delete ctx;
    return S355_OK;
  });
}

int s355_ctx_sizes(const s355_ctx* ctx, s355_sizes* out) {
  return guarded([&] {
    if (!ctx || !out) throw std::invalid_argument("null argument");
    sizes_of(ctx->dev->mesh, out);
    return S355_OK;
  });
}

int s355_ctx_set_params(s355_ctx* ctx, const s355_params* params) {
  return guarded([&] {
    check_params(params);
    dev(ctx).set_params(*params);
    ctx->ns->set_params(*params);
    return S355_OK;
  });
}

int s355_get_mesh_int(const s355_ctx* ctx, int which, int* out, size_t n) {
  return guarded([&] {
    if (!ctx) throw std::invalid_argument("null context");
    return mesh_int(ctx->dev->mesh, which, out, n);
  });
}

int s355_get_mesh_dbl(const s355_ctx* ctx, int which, double* out, size_t n) {
  return guarded([&] {
    if (!ctx) throw std::invalid_argument("null context");
    return mesh_dbl(ctx->dev->mesh, which, out, n);
  });
}

int s355_set_field(s355_ctx* ctx, int field, const double* host, size_t n) {
  return guarded([&] {
    s355::Context& d = dev(ctx);
    s355::FieldRef f = d.field(field);
    if (!host || n != f.n) throw std::invalid_argument("set_field: expected length " + std::to_string(f.n));
    S355_HIP(hipMemcpyAsync(f.p, host, n * sizeof(double), hipMemcpyHostToDevice, d.stream));
    d.sync();
    return S355_OK;
  });
}

int s355_get_field(s355_ctx* ctx, int field, double* host, size_t n) {
  return guarded([&] {
    s355::Context& d = dev(ctx);
    s355::FieldRef f = d.field(field);
    if (!host || n != f.n) throw std::invalid_argument("get_field: expected length " + std::to_string(f.n));
    S355_HIP(hipMemcpyAsync(host, f.p, n * sizeof(double), hipMemcpyDeviceToHost, d.stream));
    d.sync();
    return S355_OK;
  });
}

int s355_init_flow_field(s355_ctx* ctx) { return guarded([&] { dev(ctx).init_flow_field(); return S355_OK; }); }
int s355_synchronize(s355_ctx* ctx) {
  return guarded([&] {
    dev(ctx).sync();
    dev(ctx).check_flags();
    return S355_OK;
  });
}

int s355_bc_velocity(s355_ctx* ctx) { return guarded([&] { dev(ctx).bc_velocity(); return S355_OK; }); }
int s355_bc_pressure(s355_ctx* ctx) { return guarded([&] { dev(ctx).bc_pressure(); return S355_OK; }); }
int s355_bc_deltap(s355_ctx* ctx) { return guarded([&] { dev(ctx).bc_deltap(); return S355_OK; }); }

int s355_gradient(s355_ctx* ctx, int field, int bfield, int out) {
  return guarded([&] { dev(ctx).gradient(field, bfield, out); return S355_OK; });
}

int s355_gradient_noalloc(s355_ctx* ctx, int field, int bfield, int out) {
  return guarded([&] { dev(ctx).gradient_noalloc(field, bfield, out); return S355_OK; });
}

int s355_build_velocity_matrix(s355_ctx* ctx) {
  return guarded([&] { dev(ctx).build_velocity_matrix(); return S355_OK; });
}

int s355_solve_velocity(s355_ctx* ctx, double rtol, int maxit, s355_solve_info info[3]) {
  return guarded([&] {
    if (!info) throw std::invalid_argument("null info");
    const int rc = dev(ctx).solve_velocity(rtol, maxit, info);
    for (int k = 0; k < 3 && rc == S355_NOT_CONVERGED; ++k)
      if (info[k].reason < 0) note_not_converged("UVW" + k, 1, info[k]);
    return rc;
  });
}

int s355_cal_ru_face(s355_ctx* ctx) { return guarded([&] { dev(ctx).cal_ru_face(); return S355_OK; }); }
int s355_build_pressure_matrix(s355_ctx* ctx) {
  return guarded([&] { dev(ctx).build_pressure_matrix(); return S355_OK; });
}

int s355_solve_pressure(s355_ctx* ctx, double rtol, int maxit, s355_solve_info* info) {
  return guarded([&] {
    if (!info) throw std::invalid_argument("null info");
    const int rc = dev(ctx).solve_pressure(rtol, maxit, info);
    if (rc == S355_NOT_CONVERGED) note_not_converged("deltaP", 6, *info);
    return rc;
  });
}

int s355_correct_fields(s355_ctx* ctx, double* res_p) {
  return guarded([&] {
    s355::Context& d = dev(ctx);
    if (res_p) {
      double before[10], after[10];
      d.read_residuals(before);
      d.correct_fields();
      d.read_residuals(after);
      *res_p = after[3] - before[3];
    } else {
      d.correct_fields();
    }
    return S355_OK;
  });
}

int s355_correct_ru_face(s355_ctx* ctx) { return guarded([&] { dev(ctx).correct_ru_face(); return S355_OK; }); }
int s355_save_velocity(s355_ctx* ctx) { return guarded([&] { dev(ctx).save_velocity(); return S355_OK; }); }
int s355_velocity_residual(s355_ctx* ctx) {
  return guarded([&] { dev(ctx).velocity_residual(); return S355_OK; });
}
int s355_run_phase(s355_ctx* ctx, int phase) {
  return guarded([&] { dev(ctx).run_phase(phase); return S355_OK; });
}

int s355_residuals(s355_ctx* ctx, double res[10], int reset) {
  return guarded([&] {
    if (!res) throw std::invalid_argument("null res");
    s355::Context& d = dev(ctx);
    d.read_residuals(res);
    d.check_flags();
    if (reset) d.zero_residuals();
    return S355_OK;
  });
}

int s355_kepsilon_sources(s355_ctx* ctx) { return guarded([&] { dev(ctx).kepsilon_sources(); return S355_OK; }); }
int s355_bc_kepsilon(s355_ctx* ctx) { return guarded([&] { dev(ctx).bc_kepsilon(); return S355_OK; }); }
int s355_build_scalar_matrix(s355_ctx* ctx, int isca) {
  return guarded([&] { dev(ctx).build_scalar_matrix(isca); return S355_OK; });
}
int s355_eps_wall_rows(s355_ctx* ctx) { return guarded([&] { dev(ctx).eps_wall_rows(); return S355_OK; }); }
int s355_solve_scalar(s355_ctx* ctx, int isca, double rtol, int maxit, s355_solve_info* info) {
  return guarded([&] {
    if (!info) throw std::invalid_argument("null info");
    const int rc = dev(ctx).solve_scalar(isca, rtol, maxit, info);
    if (rc == S355_NOT_CONVERGED) note_not_converged(isca == 1 ? "Tn" : (isca == 2 ? "TE" : "ED"), 2, *info);
    return rc;
  });
}
int s355_kepsilon_post_k(s355_ctx* ctx) { return guarded([&] { dev(ctx).kepsilon_post_k(); return S355_OK; }); }
int s355_kepsilon_post_eps(s355_ctx* ctx) { return guarded([&] { dev(ctx).kepsilon_post_eps(); return S355_OK; }); }

int s355_log_stage_push(s355_ctx* ctx, int stage) { return guarded([&] { dev(ctx).stage_push(stage); return S355_OK; }); }
int s355_log_stage_pop(s355_ctx* ctx) { return guarded([&] { dev(ctx).stage_pop(); return S355_OK; }); }
int s355_log_stage_times(s355_ctx* ctx, double ms[6], int reset) {
  return guarded([&] {
    if (!ms) throw std::invalid_argument("null ms");
    dev(ctx).stage_times(ms, reset != 0);
    return S355_OK;
  });
}

int s355_calculate_velocity(s355_ctx* ctx, s355_solve_info info[3]) {
  return guarded([&] {
    dev(ctx);
    ctx->ns->CalculateVelocity();
    if (info)
      for (int k = 0; k < 3; ++k) info[k] = ctx->ns->last[k];
    return S355_OK;
  });
}

int s355_calculate_pressure(s355_ctx* ctx, s355_solve_info* info) {
  return guarded([&] {
    dev(ctx);
    ctx->ns->CalculatePressure();
    if (info) *info = ctx->ns->last[3];
    return S355_OK;
  });
}

int s355_update_tur_kepsilon(s355_ctx* ctx, s355_solve_info info[2]) {
  return guarded([&] {
    dev(ctx);
    ctx->ns->UpdateTurKEpsilon();
    if (info) {
      info[0] = ctx->ns->last[4];
      info[1] = ctx->ns->last[5];
    }
    return S355_OK;
  });
}

int s355_energy_coefficients(s355_ctx* ctx) {
  return guarded([&] {
    dev(ctx).energy_coefficients();
    return S355_OK;
  });
}

int s355_save_transient_old_data(s355_ctx* ctx) {
  return guarded([&] {
    dev(ctx).save_old_data();
    return S355_OK;
  });
}

int s355_unsteady_sources(s355_ctx* ctx, int isca) {
  return guarded([&] {
    dev(ctx).unsteady_sources(isca);
    return S355_OK;
  });
}

int s355_time_state(s355_ctx* ctx, double* cur_time, int* time_steps) {
  return guarded([&] {
    dev(ctx);
    if (!cur_time || !time_steps) throw std::invalid_argument("null output");
    *cur_time = ctx->ns->cur_time;
    *time_steps = ctx->ns->time_steps();
    return S355_OK;
  });
}

int s355_bc_temperature(s355_ctx* ctx) {
  return guarded([&] {
    dev(ctx).bc_temperature();
    return S355_OK;
  });
}

int s355_update_energy(s355_ctx* ctx, s355_solve_info* info) {
  return guarded([&] {
    dev(ctx);
    ctx->ns->UpdateEnergy();
    if (info) *info = ctx->ns->last[6];
    return S355_OK;
  });
}

int s355_last_solve_info(s355_ctx* ctx, s355_solve_info info[7]) {
  return guarded([&] {
    dev(ctx);
    if (!info) throw std::invalid_argument("null info");
    for (int k = 0; k < 7; ++k) info[k] = ctx->ns->last[k];
    return S355_OK;
  });
}

int s355_simple_iterate(s355_ctx* ctx, int n_steps, double* history, int* steps_done) {
  int done = 0;
  int rc = guarded([&] {
    dev(ctx);
    if (n_steps < 0) throw std::invalid_argument("n_steps < 0");
    return ctx->ns->NSSolve(n_steps, history, &done);
  });
  if (steps_done) *steps_done = done;
  return rc;
}

int s355_spmv_time(s355_ctx* ctx, int which, int precision, int repeats, double* ms_per_call) {
  return guarded([&] {
    if (!ms_per_call || which < 0 || which > 2 || precision < 0 || precision > 1)
      throw std::invalid_argument("spmv_time arguments");
    *ms_per_call = dev(ctx).time_spmv(which, precision, repeats);
    return S355_OK;
  });
}

int s355_kernel_shapes(s355_ctx* ctx, int lanes[2], int nt[2], int rows[2], int sigma[2], int tuned[2],
                       int slices[2], int order[2]) {
  return guarded([&] {
    if (!lanes || !nt || !rows || !sigma || !tuned || !slices || !order) throw std::invalid_argument("null output");
    dev(ctx).kernel_shapes(lanes, nt, rows, sigma, tuned, slices, order);
    return S355_OK;
  });
}

int s355_spmv_profile(s355_ctx* ctx, int precision, int mode, double* avg_ms, int* count) {
  return guarded([&] {
    if (!avg_ms || !count) throw std::invalid_argument("null output");
    *avg_ms = dev(ctx).spmv_profile(precision, mode, count);
    return S355_OK;
  });
}

int s355_set_kernel_shapes(s355_ctx* ctx, int precision, int lanes, int nt, int rows, int sigma) {
  return guarded([&] {
    dev(ctx).set_kernel_shapes(precision, lanes, nt, rows, sigma);
    return S355_OK;
  });
}

int s355_tune_report(s355_ctx* ctx, int precision, double spmv_ms[16], double gs_ms[5], int* gs_it) {
  return guarded([&] {
    if (!spmv_ms || !gs_ms || !gs_it || precision < 0 || precision > 1) throw std::invalid_argument("tune_report");
    const s355::TuneReport& r = dev(ctx).tune_report(precision);
    for (int k = 0; k < 16; ++k) spmv_ms[k] = r.spmv_ms[k];
    for (int k = 0; k < 5; ++k) gs_ms[k] = r.gs_ms[k];
    *gs_it = r.gs_it;
    return S355_OK;
  });
}

int s355_copy_time(s355_ctx* ctx, size_t bytes, int repeats, double* ms_per_call) {
  return guarded([&] {
    if (!ms_per_call) throw std::invalid_argument("null output");
    *ms_per_call = dev(ctx).time_copy(bytes, repeats);
    return S355_OK;
  });
}

}  // extern "C"
