// Batched uneven transforms in one persistent launch (kernels/fft_batch.hip).
//
// A batch is a list of independent N-D transforms (each its own extents,
// all in one precision) packed back to back in one buffer.  Every axis of
// every transform becomes one LDS pass; all passes are flattened into a
// level-major work list run by one persistent grid, captured together with
// its counter-clearing kernel into a HIP graph per direction.  The tile
// shape (workgroup size, grid oversubscription) is picked by timing the
// candidate schedules on the device (autotune) or taken from defaults.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <string>
#include <vector>

#include "planner.hpp"

namespace sa {

struct BatchSchedule {
  int block = 512;    // threads per workgroup (tile = block / T lines)
  int oversub = 1;    // persistent workgroups = oversub * resident capacity
  bool levels = true; // one launch per pass level instead of dependency waits
};

// Re-tile pass descriptor `p` for the persistent batch kernel (generic LDS
// path, padded rows, 16 elements per thread, no register stages): tiles of
// C lines of T = ceil(n / 16) threads within `block` threads and `lds_budget`
// bytes.  Returns false when one line does not fit.
bool batch_retile(sa_pass_t& p, int elem_bytes, int block, int lds_budget);

// LDS bytes the batch kernel may use per workgroup (159 KiB: its own queue
// slot sits in static LDS).
constexpr int kBatchLdsBudget = 159 * 1024;

// Work queues of an independent (level-mode) batch launch of `nitems`
// items: 8, one per XCD, for launches of at least 2048 items; 1 for smaller
// ones (a workgroup that runs out of items tries every queue before it
// exits, which costs more than the L2 locality returns on a short launch:
// micro-batches of the serving stream's small jobs 137k -> 82k jobs/s) and
// always with SHAFFT_AMD_BATCH_QUEUES=1 (read once).
uint32_t batch_level_queues(size_t nitems);

class BatchPlan {
 public:
  BatchPlan(int count, const int* ndims, const int* dims_flat, int prec);
  ~BatchPlan();
  int count() const { return (int)tr_.size(); }
  int64_t total_elems() const { return total_; }
  int64_t offset(int t) const { return tr_[t].off; }
  int64_t elems(int t) const { return tr_[t].elems; }
  int prec() const { return prec_; }
  // Attach buffers (each total_elems complex elements); in may equal out.
  void set_buffers(void* in, void* out, void* work);
  void set_stream(hipStream_t s) { stream_ = s; }
  hipStream_t stream() const { return stream_; }
  // dir -1 forward / +1 backward; graph replay unless use_graph is false.
  void execute(int dir, bool use_graph);
  // Wait for the stream; fails if any launch since the last check reported
  // a timed-out dependency wait (the flag is sticky until read).
  void synchronize();
  // Radix-8/16 stages: 0 FP32 (default), 1 bfloat16 hi/lo split on the
  // matrix cores, 2 float16 hi/lo split (complex float batches only).
  void set_precision_mode(int mode);
  int precision_mode() const { return mfma_; }
  // Time candidate schedules on the attached buffers and keep the fastest.
  std::string autotune();
  const BatchSchedule& schedule() const { return sched_; }
  int64_t total_bytes() const;  // algorithmic HBM bytes per direction
  int npasses() const { return (int)prog_.npasses(); }
  int64_t nitems() const { return (int64_t)items_.size(); }
  int grid() const { return grid_; }

 private:
  struct Tr {
    int64_t off, elems;
    int first, np;
  };
  int prec_;
  PassProgram prog_;
  std::vector<Tr> tr_;
  int64_t total_ = 0;
  void *in_ = nullptr, *out_ = nullptr, *work_ = nullptr;
  hipStream_t stream_ = nullptr;
  BatchSchedule sched_;
  int grid_ = 0, block_ = 0;
  size_t lds_ = 0;
  std::vector<sa_batch_item_t> items_;
  std::vector<size_t> level_start_;  // items_ index where each level begins (+ end)
  DevBuf d_pass_[2], d_tiles_, d_dep_, d_sign_[2], d_done_, d_head_, d_items_;
  hipGraphExec_t graph_[2] = {nullptr, nullptr};
  hipGraph_t gdef_[2] = {nullptr, nullptr};
  bool built_ = false;
  int mfma_ = 0;

  void build();  // descriptors, work list and device arrays for sched_
  void drop_graphs();
  void launch(int dir, hipStream_t s);
  void copy_identities(hipStream_t s);  // transforms with no pass
  void check_timeout_flag();            // read and clear the device flag
};

}  // namespace sa
