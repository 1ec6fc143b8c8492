#include "export.hpp"

#include <stdexcept>

#include "../device/common.hpp"
#include "../device/context.hpp"

namespace s355host {

namespace {
constexpr int kArrays = 10;  // R U V W P T VisLam VisTur TE ED
}

FieldExport::~FieldExport() {
  if (writer_.joinable()) writer_.join();
  if (done_) (void)hipEventDestroy(done_);
  if (pinned_) (void)hipHostFree(pinned_);
}

void FieldExport::wait() {
  if (writer_.joinable()) writer_.join();
  if (err_) {
    std::exception_ptr e = err_;
    err_ = nullptr;
    std::rethrow_exception(e);
  }
}

void FieldExport::snapshot() {
  wait();  // the staging buffer may still be read by the previous writer
  const size_t n = dev_.dm.Nc;
  if (!pinned_ || n_ != n) {
    if (pinned_) S355_HIP(hipHostFree(pinned_));
    pinned_ = nullptr;
    S355_HIP(hipHostMalloc(reinterpret_cast<void**>(&pinned_), kArrays * n * sizeof(double), hipHostMallocDefault));
    n_ = n;
  }
  if (!done_) S355_HIP(hipEventCreateWithFlags(&done_, hipEventDisableTiming));
  const s355::DFields& f = dev_.df;
  const double* src[kArrays] = {f.R, f.U, f.V, f.W, f.P, f.T, f.mu, f.mut, f.TE, f.ED};
  for (int k = 0; k < kArrays; ++k)
    S355_HIP(hipMemcpyAsync(pinned_ + k * n, src[k], n * sizeof(double), hipMemcpyDeviceToHost, dev_.stream));
  S355_HIP(hipEventRecord(done_, dev_.stream));
  have_ = true;
}

ResultView FieldExport::view() const {
  ResultView r;
  r.n_vert = dev_.mesh.n_vert;
  r.n_cell = dev_.mesh.n_cell;
  r.vert = dev_.mesh.vert.data();
  r.cell_vert = dev_.mesh.c_vert.data();
  r.turbulent = dev_.par.turb_model == 1;
  if (!pinned_) return r;  // mesh only
  const double* a = pinned_;
  const size_t n = n_;
  r.R = a;
  r.U = a + n;
  r.V = a + 2 * n;
  r.W = a + 3 * n;
  r.P = a + 4 * n;
  r.T = a + 5 * n;
  r.vislam = a + 6 * n;
  r.vistur = a + 7 * n;
  r.te = a + 8 * n;
  r.ed = a + 9 * n;
  return r;
}

void FieldExport::write(const std::vector<std::pair<OutFormat, std::string>>& files, bool async) {
  bool fields = false;
  for (const auto& f : files) fields |= f.first != OutFormat::GridBinary && f.first != OutFormat::GridAscii;
  if (fields && !have_) throw std::invalid_argument("output: no field snapshot taken");
  const ResultView r = view();
  hipEvent_t ev = have_ ? done_ : nullptr;
  auto job = [this, r, ev, files] {
    if (ev) S355_HIP(hipEventSynchronize(ev));
    for (const auto& f : files) write_result(r, f.first, f.second);
  };
  if (!async) {
    wait();
    job();
    return;
  }
  // a writer still running goes first: the new thread joins it, so files
  // are written in the order asked and one error is kept
  std::thread prev = std::move(writer_);
  writer_ = std::thread([this, job, prev = std::move(prev)]() mutable {
    if (prev.joinable()) prev.join();
    try {
      job();
    } catch (...) {
      if (!err_) err_ = std::current_exception();
    }
  });
}

}  // namespace s355host
