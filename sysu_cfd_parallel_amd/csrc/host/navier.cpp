// Host NavierStokesSolver: the reference's SIMPLE driver logic over the C ABI.
//   CalculateVelocity   navier_velocity.cpp:10-60
//   CalculatePressure   navier_pressure.cpp:9-97
//   UpdateTurKEpsilon   navier_scalar.cpp:11-197
//   UpdateEnergy        navier_scalar.cpp:204-326
//   NSSolve             navier.cpp:52-141 (steady: one SIMPLE pass per step)
// Log stages follow navier.cpp:26-37 (1 BuildVelocity, 2 SolveVelocity,
// 3 BuildPressure, 4 SolvePressure, 5 other).  A failed solve throws
// ConvergeError as the reference does; the caller turns it into a return
// code instead of the reference's blocking abort (BasicType.h:23-55).
#include "navier.hpp"

#include <algorithm>
#include <cstdio>
#include <filesystem>

namespace s355host {

namespace {
void throw_if(int rc, const char* what) {
  if (rc < 0) throw DeviceError{rc, std::string(what) + ": " + s355_last_error()};
}
void check_solve(int rc, const s355_solve_info& info, const char* var) {
  throw_if(rc, var);
  if (rc == S355_NOT_CONVERGED || info.reason < 0) throw ConvergeError{info.iters, info.rnorm, var, info.reason};
}
}  // namespace

void NavierStokesSolver::check(int rc, const char* what) { throw_if(rc, what); }

void DataPartition::solveVelocity_GMRES(double tol, int maxIter, s355_solve_info info[3]) {
  int rc = s355_solve_velocity(ctx_, tol, maxIter, info);
  throw_if(rc, "solveVelocity_GMRES");
  static const char* names[3] = {"Un", "Vn", "Wn"};
  for (int k = 0; k < 3; ++k)
    if (info[k].reason < 0) throw ConvergeError{info[k].iters, info[k].rnorm, names[k], info[k].reason};
}

void DataPartition::solvePressureCorrection(double tol, int maxIter, s355_solve_info* info) {
  int rc = s355_solve_pressure(ctx_, tol, maxIter, info);
  check_solve(rc, *info, "Pn");
}

void DataPartition::solveScarlar_GMRES(double tol, int maxIter, int isca, s355_solve_info* info) {
  int rc = s355_solve_scalar(ctx_, isca, tol, maxIter, info);
  check_solve(rc, *info, isca == 1 ? "Tn" : (isca == 2 ? "TE" : "ED"));
}

int NavierStokesSolver::Gradient(int phi, int bphi, int phigd) {
  check(s355_gradient(ctx_, phi, bphi, phigd), "Gradient");
  return 0;
}
void NavierStokesSolver::SetBCVelocity() { check(s355_bc_velocity(ctx_), "SetBCVelocity"); }
void NavierStokesSolver::SetBCPressure() { check(s355_bc_pressure(ctx_), "SetBCPressure"); }
void NavierStokesSolver::SetBCDeltaP() { check(s355_bc_deltap(ctx_), "SetBCDeltaP"); }
void NavierStokesSolver::SetBCKEpsilon() { check(s355_bc_kepsilon(ctx_), "SetBCKEpsilon"); }

// BC, four fused gradients and the cell loop all happen behind this entry
// (navier_velocity.cpp:74-79 then 100-383).
void NavierStokesSolver::BuildVelocityMatrix() { check(s355_build_velocity_matrix(ctx_), "BuildVelocityMatrix"); }
void NavierStokesSolver::CalRUFace2() { check(s355_cal_ru_face(ctx_), "CalRUFace2"); }
void NavierStokesSolver::BuildPressureMatrix() { check(s355_build_pressure_matrix(ctx_), "BuildPressureMatrix"); }
void NavierStokesSolver::CorrectRUFace2() { check(s355_correct_ru_face(ctx_), "CorrectRUFace2"); }
void NavierStokesSolver::BuildScalarMatrix(int iSca) {
  check(s355_build_scalar_matrix(ctx_, iSca), "BuildScalarMatrix");
}

// With params.graphs the device work between the solves runs as captured
// phases (s355_run_phase): the same member calls in the same order.
int NavierStokesSolver::CalculateVelocity() {
  check(s355_log_stage_push(ctx_, 1), "log");
  if (par_.graphs) {
    check(s355_run_phase(ctx_, S355_PHASE_MOMENTUM), "BuildVelocityMatrix");
  } else {
    BuildVelocityMatrix();
    check(s355_save_velocity(ctx_), "save velocity");
  }
  check(s355_log_stage_pop(ctx_), "log");
  check(s355_log_stage_push(ctx_, 2), "log");
  const s355_ksp_opts& o = par_.ksp[S355_KSP_VELOCITY];
  try {
    dataPartition.solveVelocity_GMRES(o.rtol, o.maxit, &last[0]);
  } catch (...) {
    s355_log_stage_pop(ctx_);
    throw;
  }
  check(s355_log_stage_pop(ctx_), "log");
  check(s355_log_stage_push(ctx_, 1), "log");
  if (par_.graphs)
    check(s355_run_phase(ctx_, S355_PHASE_VELOCITY_RESIDUAL), "velocity residual");
  else
    check(s355_velocity_residual(ctx_), "velocity residual");
  check(s355_log_stage_pop(ctx_), "log");
  return 0;
}

int NavierStokesSolver::CalculatePressure() {
  check(s355_log_stage_push(ctx_, 3), "log");
  if (par_.graphs) {
    check(s355_run_phase(ctx_, S355_PHASE_PRESSURE), "BuildPressureMatrix");
  } else {
    SetBCVelocity();
    CalRUFace2();
    BuildPressureMatrix();
  }
  check(s355_log_stage_pop(ctx_), "log");
  check(s355_log_stage_push(ctx_, 4), "log");
  const s355_ksp_opts& o = par_.ksp[S355_KSP_PRESSURE];
  try {
    dataPartition.solvePressureCorrection(o.rtol, o.maxit, &last[3]);
  } catch (...) {
    s355_log_stage_pop(ctx_);
    throw;
  }
  check(s355_log_stage_pop(ctx_), "log");
  check(s355_log_stage_push(ctx_, 3), "log");
  // SetBCDeltaP, Gradient(deltaP) and the correction loop (navier_pressure.cpp:42-70)
  if (par_.graphs) {
    check(s355_run_phase(ctx_, S355_PHASE_CORRECT), "correct fields");
  } else {
    check(s355_correct_fields(ctx_, nullptr), "correct fields");
    SetBCVelocity();
    CorrectRUFace2();
  }
  check(s355_log_stage_pop(ctx_), "log");
  return 0;
}

void NavierStokesSolver::UpdateTurKEpsilon() {
  check(s355_log_stage_push(ctx_, 5), "log");
  const s355_ksp_opts& o = par_.ksp[S355_KSP_SCALAR];
  try {
    if (par_.graphs) {
      check(s355_run_phase(ctx_, S355_PHASE_KEPS_K), "k-eps sources / k system");
    } else {
      check(s355_kepsilon_sources(ctx_), "k-eps sources");
      SetBCKEpsilon();
      // time terms after the boundary treatment (navier_scalar.cpp:71-90)
      if (!par_.steady) check(s355_unsteady_sources(ctx_, 2), "k-eps time terms");
      BuildScalarMatrix(2);
    }
    dataPartition.solveScarlar_GMRES(o.rtol, o.maxit, 2, &last[4]);
    if (par_.graphs) {
      check(s355_run_phase(ctx_, S355_PHASE_KEPS_EPS), "eps system");
    } else {
      check(s355_kepsilon_post_k(ctx_), "k clip / eps sources");
      BuildScalarMatrix(3);
      check(s355_eps_wall_rows(ctx_), "eps wall rows");
    }
    dataPartition.solveScarlar_GMRES(o.rtol, o.maxit, 3, &last[5]);
    if (par_.graphs)
      check(s355_run_phase(ctx_, S355_PHASE_KEPS_POST), "eps clip / mu_t");
    else
      check(s355_kepsilon_post_eps(ctx_), "eps clip / mu_t");
  } catch (...) {
    s355_log_stage_pop(ctx_);
    throw;
  }
  check(s355_log_stage_pop(ctx_), "log");
}

void NavierStokesSolver::SetBCTemperature() { check(s355_bc_temperature(ctx_), "SetBCTemperature"); }

// UpdateEnergy (navier_scalar.cpp:204-326) for DensityModel 0.
void NavierStokesSolver::UpdateEnergy() {
  check(s355_log_stage_push(ctx_, 5), "log");
  const s355_ksp_opts& o = par_.ksp[S355_KSP_SCALAR];
  try {
    check(s355_energy_coefficients(ctx_), "energy coefficients");
    SetBCTemperature();
    if (!par_.steady) check(s355_unsteady_sources(ctx_, 1), "energy time term");  // 290-304
    BuildScalarMatrix(1);
    dataPartition.solveScarlar_GMRES(o.rtol, o.maxit, 1, &last[6]);
  } catch (...) {
    s355_log_stage_pop(ctx_);
    throw;
  }
  check(s355_log_stage_pop(ctx_), "log");
}

void NavierStokesSolver::SetOutput(int noutput, bool binary, const std::string& dir) {
  check(s355_output_wait(ctx_), "output");
  noutput_ = noutput;
  output_binary_ = binary;
  outdir_ = dir;
  if (noutput_ <= 0) return;
  std::error_code ec;
  std::filesystem::create_directories(std::filesystem::path(outdir_) / "tec", ec);
  if (ec) throw DeviceError{S355_E_IO, "cannot create " + outdir_ + "/tec: " + ec.message()};
  // the grid file goes out once, at set-up (geometry.cpp:208-213)
  if (output_binary_)
    OutputGridBinary();
  else
    OutputGrid();
  // the monitor file's header (readyToSolve, navier.cpp:552-585)
  check(s355_monitor_write(ctx_, (outdir_ + "/monitorResult.mon").c_str(), cur_time, 1), "monitor file");
}

void NavierStokesSolver::OutputMoniter() {
  check(s355_monitor_write(ctx_, (outdir_ + "/monitorResult.mon").c_str(), cur_time, 0), "OutputMoniter");
}

void NavierStokesSolver::OutputGrid() {
  check(s355_output_write(ctx_, S355_OUT_GRID, (outdir_ + "/tec/grid.dat").c_str(), 0), "OutputGrid");
}

void NavierStokesSolver::OutputGridBinary() {
  check(s355_output_write(ctx_, S355_OUT_TEC_GEO, (outdir_ + "/tec/geo").c_str(), 0), "OutputGridBinary");
}

std::string NavierStokesSolver::result_path(int format) const {
  char name[32];
  switch (format) {
    case S355_OUT_TECPLOT: std::snprintf(name, sizeof name, "/tec/res%04d.dat", outputCounter); break;
    case S355_OUT_TEC_DATA: std::snprintf(name, sizeof name, "/tec/data"); break;
    default: std::snprintf(name, sizeof name, "/res.sav"); break;
  }
  return outdir_ + name;
}

void NavierStokesSolver::Output2Tecplot(bool async) {
  check(s355_output_snapshot(ctx_), "snapshot");
  check(s355_output_write(ctx_, S355_OUT_TECPLOT, result_path(S355_OUT_TECPLOT).c_str(), async), "Output2Tecplot");
}

void NavierStokesSolver::Output2TecplotBinary(bool async) {
  check(s355_output_snapshot(ctx_), "snapshot");
  check(s355_output_write(ctx_, S355_OUT_TEC_DATA, result_path(S355_OUT_TEC_DATA).c_str(), async),
        "Output2TecplotBinary");
}

void NavierStokesSolver::WriteBackupFile(bool async) {
  check(s355_output_snapshot(ctx_), "snapshot");
  check(s355_output_write(ctx_, S355_OUT_BACKUP, result_path(S355_OUT_BACKUP).c_str(), async), "WriteBackupFile");
}

void NavierStokesSolver::ReadBackupFile(const std::string& path) {
  check(s355_read_backup(ctx_, path.c_str()), "ReadBackupFile");
}

// navier.cpp:113-122: backup, then the Tecplot result, every noutput steps.
// Both come from one snapshot; the writer thread formats them while the
// next steps run on the device.
void NavierStokesSolver::output_step(int step) {
  if (noutput_ <= 0 || (step - 1) % noutput_ != 0) return;
  check(s355_output_snapshot(ctx_), "snapshot");
  const int result = output_binary_ ? S355_OUT_TEC_DATA : S355_OUT_TECPLOT;
  check(s355_output_write(ctx_, S355_OUT_BACKUP, result_path(S355_OUT_BACKUP).c_str(), 1), "WriteBackupFile");
  check(s355_output_write(ctx_, result, result_path(result).c_str(), 1), "Output2Tecplot");
}

void NavierStokesSolver::SaveTransientOldData() { check(s355_save_transient_old_data(ctx_), "SaveTransientOldData"); }

// One SIMPLE pass (navier.cpp:62-96); returns ResMax and leaves the ten
// residuals in Residual.
double NavierStokesSolver::simple_pass() {
  check(s355_residuals(ctx_, Residual, 1), "residual reset");
  CalculateVelocity();
  CalculatePressure();
  if (par_.turb_model == 1) UpdateTurKEpsilon();
  if (par_.solve_energy) UpdateEnergy();
  check(s355_residuals(ctx_, Residual, 0), "residuals");
  double resmax = 0.0;  // vec_max starts at 0 (tools.cpp:45-51)
  for (double r : Residual) resmax = std::max(resmax, r);
  return resmax;
}

// NSSolve (navier.cpp:52-141).  Steady: one SIMPLE pass per step until
// ResMax < ResidualSteady.  Unsteady: per time step SaveTransientOldData,
// cur_time += dt, then outer passes iter = 1 .. MaxOuterStep - 1 until the
// residual has dropped by more than 100 from the first pass; the run ends
// after MaxStep = (int)(total_time / dt) steps or once cur_time >= total_time.
// Calls continue the run: max_step bounds the steps of this call.
int NavierStokesSolver::NSSolve(int max_step, double* history, int* steps_done) {
  *steps_done = 0;
  const int max_time_steps = par_.steady ? 0 : static_cast<int>(par_.total_time / par_.dt);
  for (int step = 1; step <= max_step; ++step) {
    double resmax = 0.0;
    if (par_.steady) {
      cur_time += par_.dt;  // advanced in steady runs too (navier.cpp:59)
      resmax = simple_pass();
    } else {
      if (time_step_ >= max_time_steps) break;
      SaveTransientOldData();
      cur_time += par_.dt;
      ++time_step_;
      double resmax0 = 0.0;
      for (int iter = 1; iter < par_.max_outer_step; ++iter) {
        resmax = simple_pass();
        if (iter == 1) resmax0 = resmax;
        if (resmax0 / (resmax + 1.e-16) > 100.) break;
      }
    }
    if (history) {
      std::copy(Residual, Residual + 10, history + 11 * (step - 1));
      history[11 * (step - 1) + 10] = resmax;
    }
    *steps_done = step;
    output_step(par_.steady ? step : time_step_);
    if (noutput_ > 0) OutputMoniter();  // every step (navier.cpp:125)
    if (par_.steady ? resmax < par_.residual_steady : cur_time >= par_.total_time) break;
  }
  if (noutput_ > 0) {
    Output2Tecplot(false);  // the last result (navier.cpp:149), written before returning
    check(s355_output_wait(ctx_), "output");
  }
  return 0;
}

}  // namespace s355host
