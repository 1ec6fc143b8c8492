// Host-side NavierStokesSolver with the reference's member surface
// (navier.h:173-206) and DataPartition solve surface (MPIStructure.h:224-228).
// It drives the device exclusively through the C ABI in simple355.h, so the
// SIMPLE loop (navier.cpp:52-141) reads as the reference's does while every
// field, matrix and vector stays resident in HBM.
#pragma once
#include <string>

#include "../../../include/simple355.h"

namespace s355host {

// ConvergeError (BasicType.h:62-68)
struct ConvergeError {
  int iter;
  double residual;
  std::string varname;
  int reason;
};

// Device error surfaced as an exception, carrying the C ABI return code.
struct DeviceError {
  int code;
  std::string msg;
};

class DataPartition {
 public:
  explicit DataPartition(s355_ctx* c) : ctx_(c) {}
  // solveVelocity_GMRES (MPIStructure.cpp:234-375): U, V, W in place.
  void solveVelocity_GMRES(double tol, int maxIter, s355_solve_info info[3]);
  // solvePressureCorrection (MPIStructure.cpp:449-531): zero initial guess.
  void solvePressureCorrection(double tol, int maxIter, s355_solve_info* info);
  // solveScarlar_GMRES (MPIStructure.cpp:378-446): isca 1 = T, 2 = k, 3 = eps.
  void solveScarlar_GMRES(double tol, int maxIter, int isca, s355_solve_info* info);

 private:
  s355_ctx* ctx_;
};

class NavierStokesSolver {
 public:
  NavierStokesSolver(s355_ctx* ctx, const s355_params& p) : ctx_(ctx), par_(p), dataPartition(ctx) {}

  void set_params(const s355_params& p) { par_ = p; }

  // SIMPLE phases (navier.cpp:68, 70, 78)
  int CalculateVelocity();
  int CalculatePressure();
  void UpdateTurKEpsilon();
  void UpdateEnergy();

  // members they are made of
  void BuildVelocityMatrix();
  void CalRUFace2();
  void BuildPressureMatrix();
  void CorrectRUFace2();
  void BuildScalarMatrix(int iSca);
  int Gradient(int phi, int bphi, int phigd);
  void SetBCVelocity();
  void SetBCPressure();
  void SetBCDeltaP();
  void SetBCKEpsilon();
  void SetBCTemperature();

  // NSSolve loop; history gets 11 doubles per step (time step when
  // unsteady: the residuals of its last outer pass).
  int NSSolve(int max_step, double* history, int* steps_done);
  void SaveTransientOldData();
  double cur_time = 0.0;  // physical time of an unsteady run
  int time_steps() const { return time_step_; }

  // result and restart files (navierIO.cpp), under the output directory
  void SetOutput(int noutput, bool binary, const std::string& dir);
  void Output2Tecplot(bool async);
  void Output2TecplotBinary(bool async);
  void OutputGrid();
  void OutputGridBinary();
  void WriteBackupFile(bool async);
  void OutputMoniter();
  void ReadBackupFile(const std::string& path);

  double Residual[10] = {0};
  s355_solve_info last[7] = {};  // u, v, w, p, k, eps, T

 private:
  void check(int rc, const char* what);
  void output_step(int step);
  double simple_pass();
  int time_step_ = 0;  // time steps taken (unsteady)
  std::string result_path(int format) const;
  s355_ctx* ctx_;
  s355_params par_;
  int noutput_ = 0;  // 0: no output (the reference's default is 1, navier.cpp:515)
  bool output_binary_ = false;
  std::string outdir_;
  int outputCounter = 0;  // never advanced by the reference (navier.cpp:750)

 public:
  DataPartition dataPartition;
};

}  // namespace s355host
