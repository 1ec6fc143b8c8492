// Device context: HBM-resident mesh, fields and matrices of one case, and
// the device-level operations behind the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <memory>
#include <string>
#include <vector>

#include "../mesh/ingest.hpp"
#include "fv.hpp"
#include "krylov.hpp"

namespace s355 {

struct FieldRef {
  double* p = nullptr;
  size_t n = 0;
};

class Context {
 public:
  // ncase > 1: independent cases concatenated into one batch (see
  // concat_meshes); every operation runs all of them in one launch per kernel
  Context(const s355_mesh_desc* descs, int ncase, const s355_params& params);
  ~Context();
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;

  Mesh mesh;  // host copy (reference numbering; batch layout for ncase > 1)
  s355_params par;
  int ncase = 1;
  BatchLayout lay;  // case offsets (one case: the whole mesh)
  hipStream_t stream = nullptr;
  DMesh dm{};
  DFields df{};

  FieldRef field(int id);
  void set_params(const s355_params& p);

  // --- device-level operations (reference members)
  void init_flow_field();
  void bc_velocity();
  void bc_pressure();
  void bc_deltap();
  void gradient(int field, int bfield, int out);
  void gradient_noalloc(int field, int bfield, int out);  // limiter scratch in the fixed array
  void build_velocity_matrix();
  int solve_velocity(double rtol, int maxit, s355_solve_info info[3]);
  void cal_ru_face();
  void build_pressure_matrix();
  int solve_pressure(double rtol, int maxit, s355_solve_info* info);
  void correct_fields();  // adds into residual slot 3
  void correct_ru_face();
  void save_velocity();
  void velocity_residual();  // adds into residual slots 0..2
  void kepsilon_sources();
  void bc_kepsilon();
  void build_scalar_matrix(int isca);
  void eps_wall_rows();
  int solve_scalar(int isca, double rtol, int maxit, s355_solve_info* info);
  void kepsilon_post_k();
  void kepsilon_post_eps();
  void bc_temperature();
  void energy_coefficients();
  // unsteady runs: SaveTransientOldData, and the time terms of the scalar
  // systems (isca 2: k and epsilon, after bc_kepsilon; isca 1: energy)
  void save_old_data();
  void unsteady_sources(int isca);

  // one SIMPLE phase (S355_PHASE_*): the member calls it groups, launched
  // from one captured hipGraph when params.graphs is 1
  void run_phase(int phase);

  // residual slots and error flags on the device
  void zero_residuals();
  void read_residuals(double out[10]);          // element-wise maximum over the cases
  void read_case_residuals(double* out);        // 10 per case
  // per-case records of the last solve of each kind: U, V, W, p', k, epsilon, T
  std::vector<KResult> case_info[7];
  void check_flags();  // throws on a flag raised by a kernel
  void sync();

  // device-timed log stages (PetscLogStagePush/Pop counterpart)
  void stage_push(int stage);
  void stage_pop();
  void stage_times(double ms[6], bool reset);

  // benchmark helper
  double time_spmv(int which, int precision, int repeats);
  double time_copy(size_t bytes, int repeats);
  void kernel_shapes(int lanes[2], int nt[2], int rows[2], int sigma[2], int tuned[2], int slices[2], int order[2]);
  void set_kernel_shapes(int precision, int lanes, int nt, int rows, int sigma);
  // mode 1: start timing the Arnoldi SpMV launches of one precision (reset),
  // 0: read, -1: read and stop; returns the mean ms and the launches timed
  double spmv_profile(int precision, int mode, int* count);
  const TuneReport& tune_report(int precision) const { return precision ? kwork32_.report : kwork_.report; }

 private:
  template <class T>
  T* alloc(size_t n);
  void upload_mesh();
  std::vector<int> slice_cases(const std::vector<int>& ptr, const std::vector<int>& srow) const;
  void compute_geometry();
  void allocate_fields();
  void free_all();
  int run_solve(int which_matrix, const double* b, double* x, const s355_ksp_opts& o, bool zero_guess,
                s355_solve_info* info, int slot);
  Tiles make_tiles(const std::vector<int>& case_row, int tile);
  void ensure_mixed();
  void ensure_ilu();
  void ensure_sell();
  void prepare_sell(const double* csr_val, bool fp32, bool all);
  void grad(int nf, const double* const* phi, const double* const* bphi, double* const* out);
  void ensure_vertices();
  void ensure_ls();
  void ensure_transient();
  // the inner order of mixed GMRES (params.inner_order): false when it does
  // not apply or a 16-bit offset does not fit
  bool use_order(const s355_ksp_opts& o);
  // the mixed path's FP32 assembly option
  bool fp32_assembly() const { return par.precision == S355_MIXED && par.assembly_fp32 == 1; }
  void ensure_order();
  void fill_order(const double* csr_val, bool jacobi);
  KWork ordered_work(bool all);
  void tune_ordered(const CsrView<float>& Ao, int m);

  std::vector<void*> allocs_;
  double* partial_ = nullptr;
  size_t partial_len_ = 0;
  double* scalars_ = nullptr;  // [20..21] refinement norms
  double* res_ = nullptr;      // 10 residual slots per case
  double* rate_ = nullptr;     // outlet correction per case: mode, rate, m_in, m_out, a_out
  KHeader* hdr_ = nullptr;     // pinned, one Krylov header per case
  int* flags_ = nullptr;       // [0] bad boundary type, [1] inlet outflow in a scalar
  double *u0_ = nullptr, *v0_ = nullptr, *w0_ = nullptr;
  double *lim_vmax_ = nullptr, *lim_vmin_ = nullptr, *lim_tmp_ = nullptr;  // limiter scratch
  double* ls_inv_ = nullptr;  // least-squares gradients: inverse normal matrix per cell (6 doubles)
  KepsWork kw_{};
  KWork kwork_{};
  // mixed precision: FP32 matrix copy, FP32 Krylov workspace, refinement buffers
  float* a32_ = nullptr;
  KWork kwork32_{};
  IrWork ir_{};
  // inner order of mixed GMRES: rows in reverse Cuthill-McKee order
  bool have_order_ = false, order_ok_ = false;
  int order_kind_ = 0;  // params.inner_order the order was built for
  bool order_jacobi_ = false;  // a32o_ holds M^-1 A (Jacobi folded in)
  const int *ord_pos_ = nullptr, *ord_row_ = nullptr;  // natural row -> position, position -> natural row
  const int *ord_rowptr_ = nullptr, *ord_col_ = nullptr, *ord_diag_ = nullptr;  // the CSR by position
  const int* ord_src_ = nullptr;  // natural CSR position of each entry of the ordered CSR
  float* a32o_ = nullptr;         // FP32 values of the ordered CSR
  SellPat osell_[kSellVariants];  // SELL-64 copies of the ordered CSR, 16-bit column offsets
  float* osell32_[kSellVariants] = {nullptr, nullptr, nullptr};
  double* scratch_ = nullptr;  // this device's fixed scratch array (scratch.hpp)
  std::unique_ptr<GraphCache> graphs_;  // captured GMRES cycles (params.graphs)
  // captured SIMPLE phases (params.graphs): a phase is captured on its second
  // call (the first runs eagerly, so its lazy allocations happen outside a
  // capture) and replayed after; set_params drops them
  std::unique_ptr<GraphCache> phase_graphs_;
  unsigned phase_seen_ = 0;
  void phase_body(int phase);
  std::unique_ptr<SpmvTimer> timers_[2];  // in-solve SpMV timing, FP64 / FP32
  // SELL-64 copies of the matrix being solved, one per sorting window: patterns
  // and source maps built once, values filled before each GMRES solve that
  // uses them
  bool have_sell_ = false;
  double* sell64_[kSellVariants] = {nullptr, nullptr, nullptr};
  float* sell32_[kSellVariants] = {nullptr, nullptr, nullptr};
  IluLevels levels_{};
  bool have_levels_ = false;
  void* host_pinned_ = nullptr;
  bool has_outlet_ = false;
  int max_restart_ = 0;
  struct StageEvent {
    int stage;
    hipEvent_t a, b;
  };
  hipEvent_t take_event();
  void resolve_stages();
  std::vector<StageEvent> open_, done_;
  std::vector<hipEvent_t> pool_, all_events_;
  double stage_ms_[6] = {0, 0, 0, 0, 0, 0};
};

}  // namespace s355
