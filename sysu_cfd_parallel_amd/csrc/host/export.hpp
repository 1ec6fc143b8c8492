// Asynchronous export of the cell fields to the result and restart files:
// snapshot() enqueues device-to-host copies of the fields into pinned
// staging on the context stream and returns at once; write() formats a file
// from the latest snapshot, on a writer thread when asked, so the next SIMPLE
// iterations run on the device while the host writes.
#pragma once
#include <hip/hip_runtime.h>

#include <exception>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "output.hpp"

namespace s355 {
class Context;
}

namespace s355host {

class FieldExport {
 public:
  explicit FieldExport(s355::Context& dev) : dev_(dev) {}
  ~FieldExport();
  FieldExport(const FieldExport&) = delete;
  FieldExport& operator=(const FieldExport&) = delete;

  // Copies R U V W P T VisLam VisTur TE ED (stream-ordered after the work
  // already enqueued); waits for a writer still reading the previous one.
  void snapshot();
  // Writes the files from the latest snapshot; async: on a writer thread
  // that starts after the writers already running (files are written in
  // the order asked).
  void write(const std::vector<std::pair<OutFormat, std::string>>& files, bool async);
  // Joins the writer thread; rethrows its error.
  void wait();
  bool has_snapshot() const { return have_; }

 private:
  ResultView view() const;
  s355::Context& dev_;
  double* pinned_ = nullptr;  // 10 cell arrays
  size_t n_ = 0;
  hipEvent_t done_ = nullptr;
  bool have_ = false;
  std::thread writer_;
  std::exception_ptr err_;
};

}  // namespace s355host
