// Continuous-batching transform server: a request loop that admits FFTND /
// FFT1D jobs of mixed shape and precision onto a fixed set of device
// streams, as stream slots and device scratch memory free up.
//
// The reference has no counterpart (it runs one synchronous transform per
// call: shafft_internal.cpp:62-127); this layer sits beside the seam and
// reuses the same backends (NDBackend / OneDBackend) in their
// input-preserving mode, so a job's arithmetic is exactly that of a plain
// call on the same shape.
//
//  * Plans are built on planner threads (tables generated on the host in
//    long double and uploaded once) and stay resident in a cache keyed by
//    (kind, precision, dims); a plan carries no per-call state, so jobs of one
//    shape run concurrently on different streams.
//  * Scratch comes from one device arena allocated at start-up.  A job is
//    admitted when a stream slot is free and its scratch footprint fits in
//    the arena; jobs queue in arrival order and a later job that fits may
//    pass an earlier one that does not, at most `max_bypass` times.
//  * Small jobs (up to mb_max_elems elements) are micro-batched: the jobs
//    admitted in one scheduling round are grouped per precision (up to
//    mb_max_jobs a group) and a group runs as one persistent batch-kernel
//    launch per pass level on one slot (kernels/fft_batch.hip), its pass
//    descriptors and work list copied from pinned host staging in one
//    transfer.  A job's result then depends only on its own data (the batch
//    kernel's tiles never mix jobs), so it is the same alone and in any mix;
//    it equals the plain call's result within the precision's tolerance (the
//    batch kernel runs the generic LDS path where a plain call may use the
//    fixed-length kernel).  Larger jobs run their own pass launches on a slot.
//  * Completion is observed by polling per-slot events; latency is measured
//    from the caller-given arrival time to the observed completion.
#pragma once

#include <hip/hip_runtime_api.h>

#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "backend.hpp"
#include "planner.hpp"

namespace sa {

struct ServeJob {
  int kind = 0;  // 0 = FFTND over all axes (row-major dims), 1 = FFT1D of length n
  int ndim = 0;
  int dims[8] = {0};
  int64_t n = 0;  // FFT1D length
  int prec = 0;   // 0 = complex64, 1 = complex128
  int dir = 0;    // 0 = forward, 1 = backward
  int normalize = 0;  // scale the output by 1/sqrt(elements)
  int64_t elems() const;
  std::string key() const;
};

// Host-observed times of one job, steady-clock nanoseconds.
struct ServeTimes {
  int64_t arrival = 0, admit = 0, done = 0;
  int64_t gpu_ns = 0;  // device time between the job's first and last launch
};

struct ServeStats {
  int64_t submitted = 0, completed = 0, failed = 0;
  int64_t plans = 0, plan_builds = 0;
  int64_t arena_bytes = 0, arena_peak = 0, table_bytes = 0;
  double p50_ms = 0, p99_ms = 0, mean_ms = 0, max_ms = 0;
  double span_s = 0;   // first arrival -> last completion among completed jobs
  double gflop = 0;    // 5 N log2 N summed over completed jobs
  int64_t batches = 0, batched_jobs = 0;  // micro-batches formed, jobs they held
};

// First-fit allocator over one device allocation (offsets 256-byte aligned).
class Arena {
 public:
  explicit Arena(int64_t bytes);
  int64_t alloc(int64_t bytes);  // offset, or -1 when no free block fits
  void release(int64_t off, int64_t bytes);
  char* base() const { return static_cast<char*>(buf_.p); }
  int64_t size() const { return size_; }
  int64_t used() const { return used_; }
  int64_t peak() const { return peak_; }
  void reset_peak() { peak_ = used_; }
  static int64_t round(int64_t b) { return (b + 255) & ~int64_t(255); }

 private:
  DevBuf buf_;
  int64_t size_ = 0, used_ = 0, peak_ = 0;
  std::map<int64_t, int64_t> free_;  // offset -> size
};

class Server {
 public:
  struct Config {
    // Stream slots: one per hardware queue the HIP runtime gives the process
    // (GPU_MAX_HW_QUEUES, 4 by default); streams beyond that share queues in
    // pairs and the jobs on them serialise (serving stream: 4 slots 1.39-1.42x
    // the plain path, 8 slots 1.29-1.33x, 2 slots 1.31x).
    int slots = 4;
    int64_t arena_bytes = int64_t(16) << 30;
    int planners = 2;
    int max_bypass = 64;
    int64_t mb_max_elems = int64_t(1) << 20;  // micro-batch jobs up to this size (0: off)
    int mb_max_jobs = 64;                     // jobs per micro-batch
    int mb_max_inflight = 2;                  // micro-batches running per precision
  };
  explicit Server(const Config& cfg);
  ~Server();
  Server(const Server&) = delete;
  Server& operator=(const Server&) = delete;

  // Queue a job reading `in` and writing `out` (caller-owned device buffers
  // of elems() complex values each; `in` is not modified).  arrival_ns < 0
  // stamps now.  Returns the job id.
  int64_t submit(const ServeJob& job, const void* in, void* out, int64_t arrival_ns);
  // Block until job `id` finishes or timeout_ns passes (timeout_ns < 0: no
  // limit).  Returns 0 when finished (job status in *status, times in *t,
  // the job is then forgotten), 1 on timeout.
  int wait(int64_t id, int64_t timeout_ns, int* status, ServeTimes* t, std::string* msg);
  // 0 finished, 1 queued or running; unknown ids are an error.
  int query(int64_t id) const;
  void drain();  // block until nothing is queued or running
  // Micro-batching limits for jobs admitted from now on (max_elems 0: off;
  // max_inflight micro-batches of one precision run at once, so that under
  // load the next group gathers the jobs that arrive meanwhile).
  void set_batching(int64_t max_elems, int max_jobs, int max_inflight = 2);
  ServeStats stats() const;
  void reset_stats();
  static int64_t now_ns();

 private:
  enum JobState { QUEUED, RUNNING, DONE };
  struct Plan {
    enum { BUILDING, READY, FAILED } state = BUILDING;
    ServeJob desc;
    std::unique_ptr<NDBackend> nd;
    std::unique_ptr<OneDBackend> od;
    int64_t scratch[2] = {0, 0};  // bytes
    int status = 0;
    std::string msg;
    // Batch form of the input-preserving pass program: descriptors re-tiled
    // for the batch kernel (buffers unset), tiles per pass, launch shape.
    bool mb_ok = false;
    const PassProgram* prog = nullptr;
    PassProgram::SlotMap map;
    std::vector<sa_pass_t> mb_pass;
    std::vector<uint32_t> mb_tiles;
    int64_t mb_items = 0;
    int mb_block = 0;
    size_t mb_lds = 0;
  };
  struct Job {
    int64_t id = 0;
    ServeJob d;
    const Plan* plan = nullptr;
    const void* in = nullptr;
    void* out = nullptr;
    ServeTimes t;
    JobState state = QUEUED;
    int status = 0;
    std::string msg;
    int bypassed = 0;
    int slot = -1;
    int64_t off[2] = {-1, -1}, bytes[2] = {0, 0};
  };
  static constexpr int64_t kSlotFree = -1, kSlotBatch = -2;  // Slot::job markers
  struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t start = nullptr, end = nullptr;
    int64_t job = kSlotFree;  // job id, kSlotFree, or kSlotBatch (micro-batch)
    bool launching = false;
    // micro-batch on this slot: its jobs, passes and work items so far
    std::vector<int64_t> mb_jobs;
    int mb_prec = 0;
    int64_t mb_np = 0, mb_items = 0;
    void* mb_host = nullptr;  // pinned staging of descriptors and work list
    void* mb_dev = nullptr;   // device copy the batch kernel reads
  };
  static constexpr size_t kMbStageBytes = size_t(4) << 20;  // per slot

  void scheduler();
  void planner();
  int launch(Job& j, const Plan& p, Slot& s, std::string& msg);  // lock released
  int launch_mb(Slot& s, const std::vector<Job*>& js, std::string& msg);  // lock released
  bool mb_eligible(const Job& j, const Plan& p) const;
  static size_t mb_stage_bytes(int64_t np, int64_t items);
  void build_mb(Plan& p) const;  // batch form of a READY plan (planner thread)
  void finish(Job& j, int status, const std::string& msg);       // lock held
  void release(Job& j);                                          // lock held
  void destroy_slots();

  Config cfg_;
  Arena arena_;
  std::vector<Slot> slots_;
  mutable std::mutex mu_;
  std::condition_variable sched_cv_, done_cv_, plan_cv_;
  std::unordered_map<int64_t, Job> jobs_;
  std::deque<int64_t> pending_;
  std::unordered_map<std::string, std::unique_ptr<Plan>> plans_;
  std::deque<std::string> plan_queue_;
  int64_t next_id_ = 1;
  int running_ = 0;
  int device_ = 0;
  int cus_ = 256;  // compute units of the device (persistent grid sizes)
  bool stop_ = false;
  bool dirty_ = false;  // queue, plans or slots changed since the last admission scan
  std::map<std::pair<int, std::pair<int, size_t>>, int> occ_cache_;  // scheduler thread
  std::vector<std::thread> planners_;
  std::thread sched_;
  // statistics (lock held)
  int64_t n_submitted_ = 0, n_completed_ = 0, n_failed_ = 0, n_builds_ = 0;
  int64_t n_batches_ = 0, n_batched_jobs_ = 0;
  std::vector<int64_t> lat_ns_;
  int64_t first_arrival_ = INT64_MAX, last_done_ = 0;
  double gflop_ = 0;
};

}  // namespace sa
