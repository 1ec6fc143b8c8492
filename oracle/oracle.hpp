// ORACLE — test infrastructure only.
//
// A plain FP64 CPU restatement of the reference's per-iteration SIMPLE /
// k-epsilon path (Meiyim/SYSU_CFD_Parallel, NS/*.cpp), written from the
// mathematics the reference files describe, used as the parity checker of
// the HIP library and, timed, as the CPU baseline.  Nothing in the product
// links, loads or calls this code; only the tests, the smoke check and the
// benchmark's CPU-baseline leg do.
//
// Parity status of each part (what pins it) is listed in the header of the
// file that implements it.
#pragma once
#include <map>
#include <string>
#include <vector>

#include "../include/simple355.h"

namespace orc {

constexpr int kMaxFace = S355_MAX_CELL_FACES;
constexpr int kVoidBoundary = -10;   // VOID_CELL_ON_BOUNDARY (navier.h:17)
constexpr int kVoidInterface = -20;  // VOID_CELL_ON_INTERFACE (navier.h:18)
constexpr int kInnerFace = -10;      // INNER_FACE_BOUNDARY (navier.h:19)
constexpr double kSmall = 1.e-16;    // SMALL (BasicType.h:12)

struct Face {
  int bnd = 0;
  int v[4] = {0, 0, 0, 0};
  int c1 = 0, c2 = 0;
  double x[3] = {0, 0, 0}, n[3] = {0, 0, 0}, area = 0, lambda = 0, rlencos = 0;
  double xpac[3] = {0, 0, 0}, xnac[3] = {0, 0, 0};
};

struct Cell {
  int nface = 0;
  int rid = 0;
  int face[kMaxFace];
  int cell[kMaxFace];
  int v[8];
  bool poly = false;  // a quad face was split into two triangles
  double vol = 0, x[3] = {0, 0, 0};
};

struct Bnd {
  int face = 0;
  int v[4] = {0, 0, 0, 0};
  int rid = 0;
  double distance = 0;
};

struct SolveResult {
  int iters = 0;
  int reason = 0;
  double rnorm = 0, rnorm0 = 0;
};

struct Csr {
  std::vector<int> rowptr, col, diag;
  std::vector<int> slot;  // per (cell, local face): CSR position of the neighbour, -1 if boundary
  int nnz() const { return static_cast<int>(col.size()); }
};

class Oracle {
 public:
  Oracle(const s355_mesh_desc& mesh, const s355_params& params);

  // --- mesh (geometry.cpp) ---
  int Nvrt = 0, Ncel = 0, Nfac = 0, Nbnd = 0, n_warped = 0;
  std::vector<int> perm;
  std::vector<double> vert;  // 3 * Nvrt
  std::vector<Face> face;
  std::vector<Cell> cell;
  std::vector<Bnd> bnd;
  std::map<int, s355_region> region;
  std::vector<int> cell_faces_sorted;  // kMaxFace per cell, ascending face index, -1 padded
  Csr csr;
  s355_params par;

  // --- fields (navier.h:104-130) ---
  std::vector<double> R, U, V, W, P, T, TE, ED, deltaP, VisLam, VisTur, Apr;
  std::vector<double> dPdX, dUdX, dVdX, dWdX, dPhidX;  // 3 per cell
  std::vector<double> BRo, BU, BV, BW, BPre, BTE, BED, BTem, shear;
  std::vector<double> RUFace;
  std::vector<double> Au, Ap, As, bu, bv, bw, bp, bs;
  // old time levels of an unsteady run (navier.h:104-130)
  std::vector<double> Unp, Vnp, Wnp, Tnp, TEp, EDp, Unp2, Vnp2, Wnp2, Tnp2, TEp2, EDp2;
  double cur_time = 0;
  int time_step = 0;
  double localRes[10] = {0};
  SolveResult last[7];  // u, v, w, p, k, eps, T

  // --- phases (NavierStokesSolver members) ---
  void InitFlowField();
  void SetBCVelocity();
  void SetBCPressure();
  void SetBCDeltaP();
  void Gradient(const std::vector<double>& phi, const std::vector<double>& bphi, std::vector<double>& g);
  // weighted least-squares gradient (grad_scheme 1) and its per-cell inverse
  // normal matrices (6 values per cell, built once)
  void GradientLS(const std::vector<double>& phi, const std::vector<double>& bphi, std::vector<double>& g);
  void LsSetup();
  std::vector<double> ls_inv;
  void BuildVelocityMatrix();
  void CalRUFace2();
  void BuildPressureMatrix();
  void CorrectFields(double* res_p);
  void CorrectRUFace2();
  int CalculateVelocity();
  int CalculatePressure();
  int UpdateTurKEpsilon();
  void SetBCTemperature();
  int UpdateEnergy();
  int SolveVelocity(double rtol, int maxit);
  int SolvePressure(double rtol, int maxit);
  int SimpleIterate(int n_steps, double* history, int* done);
  void SaveTransientOldData();
  void Limiter_Barth(const std::vector<double>& UC, std::vector<double>& g);
  void Limiter_MLP(const std::vector<double>& UC, std::vector<double>& g);
  void Limiter_WENO(std::vector<double>& g);
  // the transient blocks of UpdateTurKEpsilon / UpdateEnergy on their arrays
  void TransientScalar(const std::vector<double>& old, const std::vector<double>& old2, std::vector<double>& ap,
                       std::vector<double>& src);

  void BuildScalarMatrix(std::vector<double>& phi, std::vector<double>& bphi, const std::vector<double>& diff,
                         const std::vector<double>& source, const std::vector<double>& app_in);
  void SetBCKEpsilon(std::vector<double>& tes, std::vector<double>& eds, std::vector<double>& apte,
                     std::vector<double>& aped, std::vector<double>& prod);
  // UpdateTurKEpsilon in its pieces, on the member work arrays below
  void KEpsilonSources();           // navier_scalar.cpp:33-64
  void KEpsilonBC();                // SetBCKEpsilon on the work arrays (line 67)
  void KEpsilonTimeTerms();         // 71-90
  void BuildKEpsilonMatrix(int isca);  // 2: k (93), 3: epsilon (125)
  void EpsWallRows();               // 127-154
  void KEpsilonPostK();             // 105-121: clip k, epsilon sources
  void KEpsilonPostEps();           // 165-178: clip epsilon, mu_t
  int SolveScalar(int isca, double rtol, int maxit);  // solveScarlar_GMRES into T / TE / ED
  // the work arrays of UpdateTurKEpsilon (navier_scalar.cpp:20-31)
  std::vector<double> ke_prod, ke_tes, ke_eds, ke_viste, ke_visted, ke_apte, ke_aped;

  std::vector<double>* field(int id);

 private:
  void ingest(const s355_mesh_desc& mesh);
  void create_faces();
  void find_face(int ic, int n1, int n2, int n3, int n4, std::vector<std::vector<int>>& node_face);
  void cell_face_info();
  void check_faces();
  void build_csr();
  void allocate();
  int type1_of(int rid) const;
  int SimplePass(double* resmax);
  void EnsureOld();
};

// Krylov solvers on a CSR matrix (krylov.cpp)
SolveResult ksp_solve(const Csr& A, const std::vector<double>& val, const std::vector<double>& b,
                      std::vector<double>& x, const s355_ksp_opts& o, bool zero_guess);

// z = M^-1 r with M the Jacobi or ILU(0) preconditioner of (A, val)
void pc_apply(const Csr& A, const std::vector<double>& val, int pc, const std::vector<double>& r,
              std::vector<double>& z);

// deterministic reductions: fixed 1024-element chunks summed in order
double chunked_dot(const double* a, const double* b, long n);

}  // namespace orc
