/*
 * simple355 — C ABI of the MI355X SIMPLE / k-epsilon hot path.
 *
 * The boundary replaces, one for one, the member surface of the reference's
 * NavierStokesSolver (NS/navier.h:173-206) and the solve surface of its
 * DataPartition (NS/MPIStructure.h:224-228).  Every entry point takes plain
 * pointers and sizes; device memory is owned by the context.  Return codes:
 *   S355_OK (0)               success
 *   S355_NOT_CONVERGED (1)    a Krylov solve hit its cap or diverged; the
 *                             s355_solve_info records say where (this is the
 *                             reference's ConvergeError, BasicType.h:62-68)
 *   < 0                       bad argument, bad mesh, HIP error, file
 *                             error; s355_last_error() holds the message.
 * Calls on one context are made from one host thread; contexts are
 * independent.  Field arrays are in the reference's cell / face / boundary
 * order (geometry.cpp:222-358), so host arrays copy in and out unpermuted.
 */
#ifndef SIMPLE355_H
#define SIMPLE355_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define S355_OK 0
#define S355_NOT_CONVERGED 1
#define S355_E_ARG (-1)
#define S355_E_HIP (-2)
#define S355_E_MESH (-3)
#define S355_E_STATE (-4)
#define S355_E_NOMEM (-5)
#define S355_E_IO (-6)

#define S355_MAX_CELL_FACES 12

/* BdRegion (BasicType.h:247-310): type1 1 wall, 2 inlet, 3 outlet,
 * 4 symmetry, 5 volume.  init = [u,v,w,rho,T,mu,k,eps] for a volume
 * (navier.cpp:621-633) and [u,v,w,p,rho,T,k,eps] for an inlet
 * (readparamfile.cpp:238-247). */
typedef struct s355_region {
  int rid;
  int type1;
  int type2;
  double fixed_value;
  double init[10];
} s355_region;

/* Element stream as RootProcess::read hands it on (RootStructure.cpp:428-487):
 * gmsh type codes 2 tri, 3 quad, 4 tet, 5 hex, 6 prism, 7 pyramid; tag =
 * physical region id; vertex ids 0-based. */
typedef struct s355_mesh_desc {
  int n_vert;
  const double* xyz; /* 3 * n_vert */
  int n_elem;
  const int* elem_type;
  const int* elem_tag;
  const int* elem_vptr; /* n_elem + 1 */
  const int* elem_vidx;
  int n_region;
  const s355_region* regions;
} s355_mesh_desc;

enum { S355_KSP_GMRES = 0, S355_KSP_CG = 1 };
enum { S355_PC_NONE = 0, S355_PC_JACOBI = 1, S355_PC_ILU0 = 2 };

/* One KSP configuration (KSPSetType / KSPGMRESSetRestart / PCSetType /
 * KSPSetTolerances, MPIStructure.cpp:248-262, 394-410, 460-481). */
typedef struct s355_ksp_opts {
  int type;
  int restart;
  int pc;
  int maxit;
  double rtol;
  double abstol;
  double dtol;
} s355_ksp_opts;

enum { S355_FP64 = 0, S355_MIXED = 1 };

enum { S355_KSP_VELOCITY = 0, S355_KSP_PRESSURE = 1, S355_KSP_SCALAR = 2 };

typedef struct s355_params {
  int turb_model;        /* TurModel: 0 laminar, 1 k-epsilon */
  int density_model;     /* DensityModel: only 0 (constant density) */
  int steady;            /* IfSteady: 1 steady (one SIMPLE pass per step),
                            0 unsteady (time steps of dt, see time_scheme) */
  int cell_pressure_ref; /* 0-based cell of the pressure reference */
  double urf[8];         /* URF[0..7] (navier.cpp:500-507) */
  double gravity[3];
  double residual_steady;
  s355_ksp_opts ksp[3]; /* velocity, pressure correction, scalars */
  int precision;        /* S355_FP64 or S355_MIXED */
  int ir_max;           /* mixed: cap on FP64 refinement steps per solve */
  double ir_inner_rtol; /* mixed: relative tolerance of each FP32 inner solve */
  int autotune;         /* 1: on the first GMRES solve of each precision, time
                           the SpMV lanes per row and the rows per thread of the
                           Gram-Schmidt kernels and keep the fastest (cached per
                           process and system shape); 0: fixed heuristics */
  int graphs;           /* 1: capture each GMRES restart cycle as a hipGraph
                           and replay it; 0: launch kernel by kernel */
  int solve_energy;     /* SolveEnergy: 1 = UpdateEnergy after the k-epsilon step */
  int inner_order;      /* mixed GMRES without ILU(0): the FP32 Arnoldi
                           cycle runs on the rows in another order, its
                           SELL-64 copies holding 16-bit column offsets
                           (when every offset fits): 1 = reverse
                           Cuthill-McKee (one case), 2 = the cells' order in
                           the input element stream (each case of a batch in
                           its own range); 0 = the natural order.  The
                           order is built on the first solve that uses it;
                           switching to the other order later is refused */
  double cp, prl, prte; /* heat capacity, laminar / turbulent Prandtl numbers
                           (navier.cpp:489-492: 1006, 0.72, 0.9) */
  int limiter;          /* gradient limiter (readparamfile.cpp:149-162): 0 none,
                           1 Barth, 2 MLP, 3 WENO (navier_gradient.cpp:51-62) */
  int time_scheme;      /* TimeScheme (unsteady): 1 Euler, 2 BDF2 */
  double dt;            /* time step (readparamfile.cpp:86-95) */
  double total_time;    /* the time loop ends once cur_time >= total_time */
  int max_outer_step;   /* MaxOuterStep: outer SIMPLE passes iter = 1 ..
                           MaxOuterStep - 1 per time step (navier.cpp:60) */
  int grad_scheme;       /* cell gradients: 0 Green-Gauss (Gradient,
                           navier_gradient.cpp:12-48), 1 weighted least squares
                           over the face neighbours (exact for linear fields
                           on any mesh; no counterpart in the reference);
                           the limiters apply to either */
  int assembly_fp32;     /* mixed precision: 1 = the momentum, pressure-
                           correction and scalar coefficients are computed and
                           accumulated in FP32 (the pressure-correction
                           diagonal and mass imbalance in FP64, so A 1 = 0
                           still holds); the FP64 refinement then solves that
                           system.  0 = FP64 assembly */
  int assembly_pad;
} s355_params;

typedef struct s355_solve_info {
  int iters;      /* Krylov iterations (all refinement steps together) */
  int reason;     /* >0 converged, <0 diverged (KSPConvergedReason codes) */
  double rnorm;   /* final residual norm used by the test */
  double rnorm0;  /* reference norm of the test */
  int refinements;/* FP64 refinement steps taken (mixed precision) */
  int pad;
} s355_solve_info;

/* Field ids for s355_set_field / s355_get_field. */
enum s355_field {
  /* per cell, n = Ncel */
  S355_F_R = 0, S355_F_U, S355_F_V, S355_F_W, S355_F_P, S355_F_T, S355_F_TE, S355_F_ED,
  S355_F_DELTAP, S355_F_VISLAM, S355_F_VISTUR, S355_F_APR,
  /* per cell, n = 3 * Ncel, [cell][xyz] */
  S355_F_DUDX, S355_F_DVDX, S355_F_DWDX, S355_F_DPDX, S355_F_DPHIDX,
  /* per boundary face, n = Nbnd */
  S355_F_BRO, S355_F_BU, S355_F_BV, S355_F_BW, S355_F_BPRE, S355_F_BTE, S355_F_BED,
  /* per boundary face, n = 3 * Nbnd */
  S355_F_SHEAR,
  /* per face, n = Nfac */
  S355_F_RUFACE,
  /* matrix values in CSR order, n = nnz */
  S355_F_AU, S355_F_AP, S355_F_AS,
  /* right-hand sides, n = Ncel */
  S355_F_BU_RHS, S355_F_BV_RHS, S355_F_BW_RHS, S355_F_BP_RHS, S355_F_BS_RHS,
  /* per boundary face, n = Nbnd: boundary temperature BTem */
  S355_F_BTEM,
  /* per cell, unsteady runs after s355_save_transient_old_data: the
   * previous time level (Unp ...) and the one before (Unp2 ...) */
  S355_F_UNP, S355_F_VNP, S355_F_WNP, S355_F_TNP, S355_F_TEP, S355_F_EDP,
  S355_F_UNP2, S355_F_VNP2, S355_F_WNP2, S355_F_TNP2, S355_F_TEP2, S355_F_EDP2,
  /* per cell, n = Ncel: the work arrays of UpdateTurKEpsilon
   * (navier_scalar.cpp:20-31): Prod, TESource, EDSource, ApTE, ApED, VisTE,
   * VisED, as the k-epsilon pieces below leave them */
  S355_F_KE_PROD, S355_F_KE_TESRC, S355_F_KE_EDSRC, S355_F_KE_APTE, S355_F_KE_APED, S355_F_KE_VISTE,
  S355_F_KE_VISED,
  S355_F_COUNT
};

/* Integer mesh arrays for s355_get_mesh_int. */
enum s355_mesh_int {
  S355_M_ELEM_PERM = 0, /* n_elem: element order after the partition sort */
  S355_M_FACE_C1,       /* Nfac */
  S355_M_FACE_C2,       /* Nfac: -10 on boundary faces */
  S355_M_FACE_BND,      /* Nfac: boundary index or -10 */
  S355_M_FACE_VERT,     /* 4 * Nfac */
  S355_M_CELL_NFACE,    /* Ncel */
  S355_M_CELL_FACE,     /* S355_MAX_CELL_FACES * Ncel, -1 padded */
  S355_M_CELL_CELL,     /* S355_MAX_CELL_FACES * Ncel, -1 padded */
  S355_M_BND_FACE,      /* Nbnd */
  S355_M_BND_RID,       /* Nbnd */
  S355_M_CELL_RID,      /* Ncel */
  S355_M_CSR_ROWPTR,    /* Ncel + 1 */
  S355_M_CSR_COL,       /* nnz */
  S355_M_CSR_DIAG,      /* Ncel */
  S355_M_CELL_VERT,     /* 8 * Ncel: vertices of each cell as a degenerate hex (CellData::vertices) */
  S355_M_INT_COUNT
};

/* Floating-point mesh metrics for s355_get_mesh_dbl. */
enum s355_mesh_dbl {
  S355_G_CELL_VOL = 0, /* Ncel */
  S355_G_CELL_X,       /* 3 * Ncel */
  S355_G_FACE_X,       /* 3 * Nfac */
  S355_G_FACE_N,       /* 3 * Nfac */
  S355_G_FACE_AREA,    /* Nfac */
  S355_G_FACE_LAMBDA,  /* Nfac */
  S355_G_FACE_RLENCOS, /* Nfac */
  S355_G_FACE_XPAC,    /* 3 * Nfac */
  S355_G_FACE_XNAC,    /* 3 * Nfac */
  S355_G_BND_DIST,     /* Nbnd */
  S355_G_VERT,         /* 3 * Nvrt: vertex coordinates after compaction */
  S355_G_COUNT
};

typedef struct s355_sizes {
  int n_vert, n_cell, n_face, n_bnd, n_interior, nnz, n_elem, n_warped;
} s355_sizes;

typedef struct s355_ctx s355_ctx;

/* Message of the last failing call on this thread. */
const char* s355_last_error(void);

/* Reference defaults: initSolverParam (navier.cpp:468-537) and the in-code
 * solve settings (navier_velocity.cpp:28, navier_pressure.cpp:28,
 * navier_scalar.cpp:98/158) on PETSc's defaults (GMRES(30), ILU(0)). */
void s355_default_params(s355_params* p);

/* Case generators (synthetic meshes in element-stream form).  kind:
 *   0 cavity  ip = {n, tets, seed}          dp = {Re, lid_u, jitter}
 *   1 channel ip = {nx,ny,nz, turbulent, seed} dp = {lx,ly,lz, jitter, split, Re, u_in, intensity}
 *   2 step    ip = {n_per_h, tets, turbulent, seed} dp = {upstream, downstream, span, jitter, Re, u_in, intensity}
 * s355_case_build returns an opaque handle; s355_case_sizes fills
 * {n_vert, n_elem, n_vidx, n_region}; s355_case_copy copies the arrays out. */
typedef struct s355_case s355_case;
int s355_case_build(int kind, const int* ip, const double* dp, s355_case** out);
int s355_case_sizes(const s355_case* c, int* sizes4);
int s355_case_copy(const s355_case* c, double* xyz, int* type, int* tag, int* vptr, int* vidx,
                   s355_region* regions);
void s355_case_free(s355_case* c);
/* Mesh files in the reference's formats: binary = 0 gmsh 2 ASCII
 * (RootProcess::read, RootStructure.cpp:428-487), 1 the binary form of
 * RootProcess::readBin (RootStructure.cpp:61-116, written by postCycas.py).
 * The first tag of each element is its region id; unreferenced vertices are
 * dropped and the rest renumbered in order (RootStructure.cpp:930-968).  A
 * case read from a file has no region table: the caller supplies one (the
 * reference's param.in). */
int s355_case_read(const char* path, int binary, s355_case** out);
/* Writes a mesh (its element stream) in either format. */
int s355_case_write(const s355_mesh_desc* mesh, const char* path, int binary);

/* --- context: set-up (ReadGridFile / CreateFaces / CellFaceInfo /
 *     CheckAndAllocate / initPetsc / InitFlowField) --------------------- */
int s355_ctx_create(const s355_mesh_desc* mesh, const s355_params* params, s355_ctx** out);
int s355_ctx_destroy(s355_ctx* ctx);
int s355_ctx_sizes(const s355_ctx* ctx, s355_sizes* out);
int s355_ctx_set_params(s355_ctx* ctx, const s355_params* params);
int s355_get_mesh_int(const s355_ctx* ctx, int which, int* out, size_t n);
int s355_get_mesh_dbl(const s355_ctx* ctx, int which, double* out, size_t n);
int s355_set_field(s355_ctx* ctx, int field, const double* host, size_t n);
int s355_get_field(s355_ctx* ctx, int field, double* host, size_t n);
/* InitFlowField (navier.cpp:588-697) */
int s355_init_flow_field(s355_ctx* ctx);
int s355_synchronize(s355_ctx* ctx);

/* --- phase entries (NavierStokesSolver members) -------------------------- */
int s355_bc_velocity(s355_ctx* ctx);                       /* SetBCVelocity  navier_bc.cpp:8-119   */
int s355_bc_pressure(s355_ctx* ctx);                       /* SetBCPressure  navier_bc.cpp:121-155 */
int s355_bc_deltap(s355_ctx* ctx);                         /* SetBCDeltaP    navier_bc.cpp:157-181 */
/* Gradient (navier_gradient.cpp:12-74): cell field `field` with boundary
 * values `bfield` into 3-vector field `out`. */
int s355_gradient(s355_ctx* ctx, int field, int bfield, int out);
/* s355_gradient for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc); no
 * counterpart in the reference.  The limiters' scratch (MLP: 2 doubles per
 * vertex, WENO: 3 per cell) lives in a fixed 24 MiB array per GPU
 * (3 * 2^20 doubles), held until the process ends; larger meshes are
 * rejected with S355_E_ARG.  The limiter's vertex arrays are uploaded when
 * the context is created or its parameters set.  Calls that overlap on one
 * GPU share the array and race. */
int s355_gradient_noalloc(s355_ctx* ctx, int field, int bfield, int out);
int s355_build_velocity_matrix(s355_ctx* ctx);             /* navier_velocity.cpp:63-393 */
int s355_solve_velocity(s355_ctx* ctx, double rtol, int maxit, s355_solve_info info[3]); /* MPIStructure.cpp:234-375 */
int s355_cal_ru_face(s355_ctx* ctx);                       /* CalRUFace2 navier_velocity.cpp:425-476 */
int s355_build_pressure_matrix(s355_ctx* ctx);             /* navier_pressure.cpp:99-232 */
int s355_solve_pressure(s355_ctx* ctx, double rtol, int maxit, s355_solve_info* info); /* MPIStructure.cpp:449-531 */
/* correction block navier_pressure.cpp:39-78 (SetBCDeltaP, Gradient(dp),
 * P/U/V/W update); adds sum |dp - dp_ref| vol into residual slot 3 and, when
 * res_p is not NULL, returns this call's value (synchronous). */
int s355_correct_fields(s355_ctx* ctx, double* res_p);
int s355_correct_ru_face(s355_ctx* ctx);                   /* CorrectRUFace2 navier_pressure.cpp:235-278 */
/* residual bookkeeping of CalculateVelocity (navier_velocity.cpp:19-24, 48-52):
 * snapshot U,V,W before the solve, then add sum |old - new| vol into
 * residual slots 0..2. */
int s355_save_velocity(s355_ctx* ctx);
int s355_velocity_residual(s355_ctx* ctx);
/* localRes -> Residual (navier.cpp:64-65, 94): read the 10 residual slots;
 * reset != 0 zeroes them afterwards. Synchronous. */
int s355_residuals(s355_ctx* ctx, double res[10], int reset);

/* The device work between the solves of one SIMPLE pass (navier.cpp:60-101)
 * in phases.  With params.graphs = 1 a phase is captured as one hipGraph on
 * its second call and replayed after (s355_ctx_set_params drops the graphs);
 * with 0 it runs kernel by kernel.  Either way it is exactly the member
 * calls listed, in that order, so results are the same bits.  No reference
 * counterpart: the reference calls the members one by one. */
enum s355_phase {
  S355_PHASE_MOMENTUM = 0,       /* build_velocity_matrix, save_velocity */
  S355_PHASE_VELOCITY_RESIDUAL,  /* velocity_residual */
  S355_PHASE_PRESSURE,           /* bc_velocity, cal_ru_face, build_pressure_matrix */
  S355_PHASE_CORRECT,            /* correct_fields (res_p NULL), bc_velocity, correct_ru_face */
  S355_PHASE_KEPS_K,             /* kepsilon_sources, bc_kepsilon, unsteady_sources(2)
                                    (unsteady runs), build_scalar_matrix(2) */
  S355_PHASE_KEPS_EPS,           /* kepsilon_post_k, build_scalar_matrix(3), eps_wall_rows */
  S355_PHASE_KEPS_POST,          /* kepsilon_post_eps */
  S355_PHASE_COUNT
};
int s355_run_phase(s355_ctx* ctx, int phase);

/* k-epsilon pieces of UpdateTurKEpsilon (navier_scalar.cpp:11-197). */
int s355_kepsilon_sources(s355_ctx* ctx);                  /* 33-64: mu_k, mu_eps, P_k, S_k, Ap_k */
int s355_bc_kepsilon(s355_ctx* ctx);                       /* SetBCKEpsilon navier_bc.cpp:220-345 */
int s355_build_scalar_matrix(s355_ctx* ctx, int isca);     /* BuildScalarMatrix 407-592; 1 = T, 2 = k, 3 = eps */
int s355_eps_wall_rows(s355_ctx* ctx);                     /* 129-154: wall-cell rows -> identity */
int s355_solve_scalar(s355_ctx* ctx, int isca, double rtol, int maxit, s355_solve_info* info); /* MPIStructure.cpp:378-446 */
int s355_kepsilon_post_k(s355_ctx* ctx);                   /* 105-121: clip k, eps sources */
int s355_kepsilon_post_eps(s355_ctx* ctx);                 /* 165-178: clip eps, mu_t */

/* Energy pieces of UpdateEnergy (navier_scalar.cpp:204-326, DensityModel 0):
 * BuildScalarMatrix(1) and solve_scalar(1) act on T / BTem. */
int s355_energy_coefficients(s355_ctx* ctx);  /* 215-218: kcond = cp (mu/prl + mu_t/prte) / cp, ESource = ApE = 0 */
int s355_bc_temperature(s355_ctx* ctx);       /* SetBCTemperature navier_bc.cpp:183-214 */

/* Unsteady runs (params.steady = 0).  The momentum time term is part of
 * s355_build_velocity_matrix (navier_velocity.cpp:107-123); with
 * params.limiter the limiter runs inside every gradient
 * (navier_gradient.cpp:51-62). */
int s355_save_transient_old_data(s355_ctx* ctx); /* SaveTransientOldData navier.cpp:700-742 */
/* time terms of the scalar systems: isca 2 adds them to the k and epsilon
 * sources / diagonals after s355_bc_kepsilon (navier_scalar.cpp:71-90),
 * isca 1 gives the energy system its source / diagonal (290-304); a no-op
 * for steady runs */
int s355_unsteady_sources(s355_ctx* ctx, int isca);
/* physical time and time steps taken by NSSolve (s355_simple_iterate) */
int s355_time_state(s355_ctx* ctx, double* cur_time, int* time_steps);

/* Device-timed log stages, the counterpart of PetscLogStagePush/Pop around
 * the phases (navier.cpp:26-37): 1 BuildVelocity, 2 SolveVelocity,
 * 3 BuildPressure, 4 SolvePressure, 5 other (turbulence, bookkeeping). */
int s355_log_stage_push(s355_ctx* ctx, int stage);
int s355_log_stage_pop(s355_ctx* ctx);
/* ms[0..5]: accumulated device time per stage (index = stage, ms[0] = total
 * of all stages); reset != 0 clears the accumulators.  Synchronous. */
int s355_log_stage_times(s355_ctx* ctx, double ms[6], int reset);

/* Phase-level entries, as NSSolve calls them (navier.cpp:68, 70, 78). */
/* These run the host NavierStokesSolver mirror owned by the context; the
 * residual contributions go to the device residual slots (s355_residuals). */
int s355_calculate_velocity(s355_ctx* ctx, s355_solve_info info[3]);
int s355_calculate_pressure(s355_ctx* ctx, s355_solve_info* info);
int s355_update_tur_kepsilon(s355_ctx* ctx, s355_solve_info info[2]);
int s355_update_energy(s355_ctx* ctx, s355_solve_info* info);  /* navier.cpp:84-86 */

/* ---- Batches of independent cases (parameter sweeps) ----------------------
 * No reference counterpart: the reference runs one case per MPI job.  The
 * cases are concatenated into one device layout (cells, rows and vertices
 * case after case; the boundary faces of all cases first, then their interior
 * faces, each case keeping its own order; region ids of case k shifted by
 * k * (largest region id + 1)), and every phase entry above runs all of them
 * with one launch per kernel.  Each case keeps its own Krylov solve, its
 * own outlet mass-flow correction, pressure reference cell (the same index
 * within every case) and residuals; s355_residuals returns the element-wise
 * maximum over the cases.  Every system is solved with GMRES. */
enum s355_batch_array {
  S355_BATCH_CELL = 0,  /* n_case + 1 offsets of the cases' cells (and CSR rows) */
  S355_BATCH_BND = 1,   /* boundary records / boundary faces */
  S355_BATCH_IFACE = 2, /* interior faces, counted from n_bnd */
  S355_BATCH_NNZ = 3,   /* CSR entries */
  S355_BATCH_VERT = 4,
  S355_BATCH_ELEM = 5
};
int s355_batch_create(int n_case, const s355_mesh_desc* meshes, const s355_params* params, s355_ctx** out);
int s355_batch_size(const s355_ctx* ctx, int* n_case);
int s355_batch_layout(const s355_ctx* ctx, int which, int* out, size_t n);
/* res: 10 per case, in case order */
int s355_batch_residuals(s355_ctx* ctx, double* res, int reset);
/* per-case records of the last solve of a kind: 0..2 U, V, W, 3 p', 4 k, 5 eps, 6 T */
int s355_batch_solve_info(s355_ctx* ctx, int kind, s355_solve_info* info, int n_case);

/* Solve records of the last SIMPLE phases run through the phase-level
 * entries: info[0..6] = U, V, W, pressure correction, k, epsilon, T. */
int s355_last_solve_info(s355_ctx* ctx, s355_solve_info info[7]);

/* NSSolve steady loop body (navier.cpp:52-141) for n_steps steps.  history
 * receives 11 doubles per step: Residual[0..9] and ResMax.  *steps_done is
 * the number of completed steps (stops early when ResMax < residual_steady
 * or a solve fails). */
int s355_simple_iterate(s355_ctx* ctx, int n_steps, double* history, int* steps_done);

/* Times the production SpMV launch y = M^-1 A x (Jacobi scaling, the kernel
 * and shape the GMRES solves use) on A in {0 Au, 1 Ap, 2 As} as last
 * assembled: precision 0 FP64, 1 FP32 copy.  *ms_per_call is the mean device
 * time of `repeats` launches, each timed alone with HIP events on the context
 * stream after a 512 MiB overwrite has cleared the Infinity Cache. */
int s355_spmv_time(s355_ctx* ctx, int which, int precision, int repeats, double* ms_per_call);

/* Mean device time of a plain device-to-device copy of `bytes` (a multiple
 * of 16) over `repeats` launches: the bandwidth reference for the roofline
 * and for scaling hardware byte counters.  No reference counterpart. */
int s355_copy_time(s355_ctx* ctx, size_t bytes, int repeats, double* ms_per_call);

/* Kernel shapes in use for the GMRES solves of each precision (index 0 FP64,
 * 1 FP32): lanes[k] SpMV lanes per row, nt[k] 1 when the SpMV loads the
 * matrix non-temporally, rows[k] rows per thread of the multi-dot /
 * Gram-Schmidt / update kernels, sigma[k] the row window the SELL-64 copy
 * is sorted in (used when lanes[k] is 0), tuned[k] 1 when they came from
 * the autotuner, slices[k] the SELL-64 slices each wave takes (1, or 2 for
 * the FP32 product on one case; 0 with CSR), order[k] 1 when the mixed
 * solves' FP32 cycle runs in the inner order (params.inner_order set and its
 * 16-bit offsets fitting; always 0 for FP64).  No reference counterpart. */
int s355_kernel_shapes(s355_ctx* ctx, int lanes[2], int nt[2], int rows[2], int sigma[2], int tuned[2],
                       int slices[2], int order[2]);

/* Times the production SpMV launches inside the GMRES solves of one
 * precision (0 FP64, 1 FP32): every 4th Arnoldi SpMV is bracketed by HIP
 * events on the context stream, without added synchronisation.  mode 1
 * starts (and resets), 0 reads, -1 reads and stops; *avg_ms is the mean
 * launch time, *count the launches timed.  Solves replayed from captured
 * graphs are not timed.  No reference counterpart. */
int s355_spmv_profile(s355_ctx* ctx, int precision, int mode, double* avg_ms, int* count);

/* Sets the kernel shapes of one precision's GMRES solves explicitly (the
 * autotuner then leaves them alone): lanes 1, 2, 4, 8, 16 (CSR SpMV) or 0
 * (SELL-64 SpMV; FP32 on one case takes two slices per wave above 2^20 rows),
 * -1 (SELL-64, one slice per wave) or -2 (SELL-64, two slices per wave; FP32
 * on one case only), nt 0/1, rows 1, 2, 4, 8, 16 or 0 (by size), sigma the
 * window of rows the SELL-64 copy sorts by length: 64, 512, 4096 or 0 (4096).
 * No reference counterpart. */
int s355_set_kernel_shapes(s355_ctx* ctx, int precision, int lanes, int nt, int rows, int sigma);

/* The autotuner's timings for one precision (0 FP64, 1 FP32): spmv_ms[v] for
 * v = 2 k + nt, k = 0..4 for CSR at 1, 2, 4, 8, 16 lanes per row, k = 5, 6, 7
 * for SELL-64 sorted in windows of 64, 512, 4096 rows; gs_ms[r] for 2^r rows
 * per thread (multi-dot, its final reduction and the Gram-Schmidt update,
 * averaged over the Arnoldi steps a quarter, half and three quarters into the
 * restart cycle; *gs_it is the middle one); -1 where not timed.  No reference
 * counterpart. */
int s355_tune_report(s355_ctx* ctx, int precision, double spmv_ms[16], double gs_ms[5], int* gs_it);

/* Host-side set-up only (no device): the integer and metric output of the
 * face builder, for inspection and for tests on machines without a GPU. */
typedef struct s355_mesh s355_mesh;
int s355_mesh_build(const s355_mesh_desc* mesh, s355_mesh** out);
/* the batch layout of s355_batch_create, host side */
int s355_mesh_build_batch(int n_case, const s355_mesh_desc* meshes, s355_mesh** out);
int s355_mesh_sizes(const s355_mesh* m, s355_sizes* out);
int s355_mesh_get_int(const s355_mesh* m, int which, int* out, size_t n);
int s355_mesh_get_dbl(const s355_mesh* m, int which, double* out, size_t n);
void s355_mesh_free(s355_mesh* m);

/* --- result and restart files (navierIO.cpp), one partition ------------ */
enum s355_output_format {
  S355_OUT_TECPLOT = 0, /* Output2Tecplot (navierIO.cpp:161-246): Tecplot ASCII, tec/resNNNN.dat */
  S355_OUT_TEC_DATA,    /* Output2TecplotBinary (navierIO.cpp:132-159): tec/data */
  S355_OUT_TEC_GEO,     /* OutputGridBinary (navierIO.cpp:29-56): tec/geo */
  S355_OUT_GRID,        /* OutputGrid (navierIO.cpp:58-122): tec/grid.dat */
  S355_OUT_BACKUP,      /* WriteBackupFile (navierIO.cpp:370-386): res.sav */
  S355_OUT_VTK,         /* Output2VTK (navierIO.cpp:287-365): res.vtk */
  S355_OUT_COUNT
};

/* Cell arrays of one result (length Ncel each); vistur, te, ed are used
 * only when turbulent (vistur may be null for a laminar result). */
typedef struct s355_result {
  const double *R, *U, *V, *W, *P, *T, *vislam, *vistur, *te, *ed;
  int turbulent;
  int pad;
} s355_result;

/* Host only: writes one file of a host mesh and host arrays (the same
 * writers the context uses).  The grid formats ignore r (may be null). */
int s355_mesh_write_result(const s355_mesh* m, int format, const char* path, const s355_result* r);
/* Host only: ReadBackupFile (navierIO.cpp:389-403) into out, laid out as
 * R U V W P T VisLam (VisTur TE ED when turbulent), n_cell values each. */
int s355_backup_read(const char* path, int n_cell, int turbulent, double* out);

/* Starts a device-to-host copy of the cell fields (R U V W P T VisLam
 * VisTur TE ED) into pinned staging, ordered after the work already
 * enqueued, and returns without waiting.  Waits first for a writer still
 * reading the previous snapshot.  No reference counterpart (the reference
 * writes from host-resident fields). */
int s355_output_snapshot(s355_ctx* ctx);
/* Writes one file in `format` from the latest snapshot (the grid formats
 * need none).  async = 1: on a writer thread, so the caller can keep the
 * device busy; errors come back from s355_output_wait. */
int s355_output_write(s355_ctx* ctx, int format, const char* path, int async);
/* Joins the writer thread; returns its error, if any. */
int s355_output_wait(s355_ctx* ctx);
/* ReadBackupFile (navierIO.cpp:389-403): sets R U V W P T VisLam (VisTur TE
 * ED with k-epsilon) from a res.sav.  InitFlowField reads the backup in
 * place of the region initial values (navier.cpp:591-592); call it after
 * s355_init_flow_field, which still sets the boundary values. */
int s355_read_backup(s355_ctx* ctx, const char* path);
/* The output schedule of NSSolve (navier.cpp:113-122, 149): every noutput
 * steps (step - 1 divisible by noutput) a backup dir/res.sav and either
 * dir/tec/res0000.dat (binary = 0) or dir/tec/data (binary = 1), written
 * from a snapshot on a writer thread while the next steps run; the grid
 * file (dir/tec/grid.dat or dir/tec/geo) now; and a last Tecplot file
 * after the loop.  noutput = 0 turns output off (the default).  The
 * reference never advances its output counter (navier.cpp:750), so every
 * ASCII result goes to res0000.dat. */
int s355_output_schedule(s355_ctx* ctx, int noutput, int binary, const char* dir);
/* OutputMoniter (navierIO.cpp:405-468): per wall region the summed wall
 * force (Bnd.shear) and per fluid region of type2 0 the summed W^2 vol, region
 * id 0 left out, ascending region id; 6 doubles per record: rid, type1,
 * fx, fy, fz, scalar. */
int s355_monitor_records(s355_ctx* ctx, double* out, int max_records, int* n_records);
/* monitorResult.mon: header = 1 starts the file with the header of
 * readyToSolve (navier.cpp:552-585); header = 0 appends the line of the
 * current records at time cur_time.  NSSolve writes it every step while an
 * output schedule is set. */
int s355_monitor_write(s355_ctx* ctx, const char* path, double cur_time, int header);

#ifdef __cplusplus
}
#endif

#endif /* SIMPLE355_H */
