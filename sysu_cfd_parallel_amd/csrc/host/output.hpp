// Result files in the reference's formats: Tecplot ASCII and the binary
// tec/geo + tec/data pair read by postCycas.py, the ASCII grid file, and the
// res.sav backup that restarts a run.  Host code only: the device side hands
// over host copies of the cell fields.
#pragma once
#include <stdexcept>
#include <string>
#include <vector>

namespace s355host {

// A file that cannot be opened, read or written, or a malformed one.
struct IoError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// Host views of one result: the vertex pool, each cell's 8 vertices
// (0-based, degenerate-hex form) and the cell fields.  te / ed / vistur may
// be null for a laminar run (turbulent == false); vistur null reads as 0.
struct ResultView {
  int n_vert = 0, n_cell = 0;
  const double* vert = nullptr;  // 3 per vertex
  const int* cell_vert = nullptr;
  const double *R = nullptr, *U = nullptr, *V = nullptr, *W = nullptr, *P = nullptr, *T = nullptr;
  const double *vislam = nullptr, *vistur = nullptr, *te = nullptr, *ed = nullptr;
  bool turbulent = false;
};

enum class OutFormat {
  TecplotAscii = 0,  // Output2Tecplot: tec/resNNNN.dat
  TecplotData = 1,   // Output2TecplotBinary: tec/data
  GridBinary = 2,    // OutputGridBinary: tec/geo
  GridAscii = 3,     // OutputGrid: tec/grid.dat
  Backup = 4,        // WriteBackupFile: res.sav
  Vtk = 5,           // Output2VTK: res.vtk (legacy ASCII unstructured grid)
};

// Writes one file; throws std::runtime_error on a bad view or an I/O error.
void write_result(const ResultView& r, OutFormat f, const std::string& path);

// Cell arrays a backup file restores (ReadBackupFile): each of length n_cell;
// vistur / te / ed are read only when turbulent.
struct BackupArrays {
  double *R = nullptr, *U = nullptr, *V = nullptr, *W = nullptr, *P = nullptr, *T = nullptr;
  double *vislam = nullptr, *vistur = nullptr, *te = nullptr, *ed = nullptr;
};
void read_backup(const std::string& path, int n_cell, bool turbulent, const BackupArrays& out);

// The monitor file (monitorResult.mon).  One record per wall region (the
// wall force: the summed Bnd.shear) and per fluid region (the summed W^2 vol
// of a type2 = 0 region), region id 0 left out, in ascending region id
// (NavierStokesSolver::readyToSolve navier.cpp:552-585 writes the header,
// OutputMoniter navierIO.cpp:405-468 a line per step).
struct MonitorRecord {
  int rid = 0;
  int type1 = 0;  // 1 wall, 5 fluid
  double force[3] = {0, 0, 0};
  double scalar = 0;
};
std::string monitor_header(const std::vector<MonitorRecord>& regions);
std::string monitor_line(double cur_time, const std::vector<MonitorRecord>& records);

}  // namespace s355host
