#include "batch.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <sstream>

#include "fastdiv.hpp"

namespace sa {

static int cu_count() {
  static const int cached = [] {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
      return v;
    return 256;
  }();
  return cached;
}

static uint64_t wall_clock_hz() {
  static const uint64_t cached = []() -> uint64_t {
    int dev = 0, khz = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) == hipSuccess && khz > 0)
      return (uint64_t)khz * 1000ull;
    return (uint64_t)100000000ull;  // gfx9 wall clock: 100 MHz
  }();
  return cached;
}

uint32_t batch_level_queues(size_t nitems) {
  static const bool one = [] {
    const char* e = std::getenv("SHAFFT_AMD_BATCH_QUEUES");
    return e && std::atoi(e) == 1;
  }();
  return (!one && nitems >= 2048) ? 8u : 1u;
}

bool batch_retile(sa_pass_t& p, int elem_bytes, int block, int lds_budget) {
  const int n = p.n;
  const int T = (n + 15) / 16;
  if (T > 512) return false;
  int C = std::max(1, block / T);
  int stride = n + (n - 1) / 16;
  if (stride % 2 == 0) stride += 1;
  while (C > 1 && (int64_t)C * stride * elem_bytes > lds_budget) --C;
  if ((int64_t)C * stride * elem_bytes > lds_budget) return false;
  if (C > p.batch) C = (int)std::max<int64_t>(1, p.batch);
  p.T = T;
  p.emax = 16;
  p.C = C;
  p.ldsstride = stride;
  p.lds_pad = 1;  // the batch kernel runs padded rows
  p.Tdiv = make_fastdiv((uint32_t)T);
  p.Cdiv = make_fastdiv((uint32_t)C);
  p.reg_first = p.reg_last = 0;
  p.nt = 0;
  p.logC = 0;
  return true;
}

BatchPlan::BatchPlan(int count, const int* ndims, const int* dims_flat, int prec)
    : prec_(prec), prog_(prec) {
  if (prec != 0 && prec != 1) fail(ERR_INVALID_FFTTYPE, "invalid FFT type");
  if (count < 1 || !ndims || !dims_flat) fail(ERR_INVALID_DIM, "empty batch");
  const int64_t cap = lds_cap(prec);
  int64_t off = 0;
  const int* d = dims_flat;
  for (int t = 0; t < count; ++t) {
    const int nd = ndims[t];
    if (nd < 1) fail(ERR_INVALID_DIM, "transform with no axes");
    int64_t elems = 1;
    for (int i = 0; i < nd; ++i) {
      if (d[i] < 1) fail(ERR_INVALID_DIM, "dimension must be positive");
      elems *= d[i];
    }
    Tr tr{off, elems, (int)prog_.npasses(), 0};
    for (int ax = nd - 1; ax >= 0; --ax) {
      const int64_t n = d[ax];
      if (n == 1) continue;
      // One LDS pass per axis: a length that needs a multi-pass split
      // would need scratch of its own and is left to the per-call plans.
      const bool fits = (n <= cap && is_smooth7(n));
      int64_t m = 1;
      while (m < 2 * n - 1) m <<= 1;
      if (!fits && m > cap) fail(ERR_NOT_IMPL, "batched axis longer than one LDS pass");
      int64_t outer = 1, inner = 1;
      for (int i = 0; i < ax; ++i) outer *= d[i];
      for (int i = ax + 1; i < nd; ++i) inner *= d[i];
      std::vector<Digit> lines = {{outer, n * inner, n * inner}, {inner, 1, 1}};
      prog_.add_transform(n, lines, inner, inner);
    }
    tr.np = (int)prog_.npasses() - tr.first;
    tr_.push_back(tr);
    off += elems;
    d += nd;
  }
  total_ = off;
}

BatchPlan::~BatchPlan() { drop_graphs(); }

void BatchPlan::drop_graphs() {
  for (int i = 0; i < 2; ++i) {
    if (graph_[i]) (void)hipGraphExecDestroy(graph_[i]);
    if (gdef_[i]) (void)hipGraphDestroy(gdef_[i]);
    graph_[i] = nullptr;
    gdef_[i] = nullptr;
  }
}

void BatchPlan::set_buffers(void* in, void* out, void* work) {
  if (!in || !out || !work) fail(ERR_NULLPTR, "batch buffers must be set");
  if (work == in || work == out) fail(ERR_INVALID_STATE, "work buffer must be distinct");
  in_ = in;
  out_ = out;
  work_ = work;
  drop_graphs();
  build();
}

void BatchPlan::build() {
  const int esz = prec_ ? 16 : 8;
  const size_t np = prog_.npasses();
  std::vector<sa_pass_t> desc[2];
  std::vector<uint32_t> tiles(np);
  std::vector<int32_t> dep(np, -1);
  std::vector<int8_t> sign[2];
  desc[0].resize(np);
  desc[1].resize(np);
  sign[0].resize(np);
  sign[1].resize(np);
  block_ = 1;
  lds_ = 0;
  const int budget = kBatchLdsBudget;
  for (const Tr& tr : tr_) {
    for (int j = 0; j < tr.np; ++j) {
      const size_t i = (size_t)(tr.first + j);
      const PassSpec& ps = prog_.pass(i);
      sa_pass_t p = ps.p;
      // Re-tile for the batch kernel (generic LDS path).
      if (!batch_retile(p, esz, sched_.block, budget)) fail(ERR_INTERNAL, "batched pass does not fit LDS");
      const int C = p.C;
      block_ = std::max(block_, (C * p.T + 63) / 64 * 64);  // whole waves
      lds_ = std::max(lds_, (size_t)C * p.ldsstride * esz);
      tiles[i] = (uint32_t)((p.batch + C - 1) / C);
      if (j > 0) dep[i] = (int32_t)(i - 1);
      // Buffers: pass j writes `out` when (np-1-j) is even, else `work`.
      auto buf_of = [&](int jj) -> char* {
        return static_cast<char*>(((tr.np - 1 - jj) % 2 == 0) ? out_ : work_);
      };
      const char* src = j == 0 ? static_cast<const char*>(in_) : buf_of(j - 1);
      char* dst = buf_of(j);
      p.in = src + tr.off * esz;
      p.out = dst + tr.off * esz;
      for (int d = 0; d < 2; ++d) {
        sa_pass_t q = p;
        if (ps.dir_tables) {
          q.pre = ps.pre_t[d];
          q.post = ps.post_t[d];
          q.blu_h = ps.blu_t[d];
        }
        desc[d][i] = q;
        sign[d][i] = (int8_t)(ps.sign == 0 ? (d == 0 ? -1 : +1) : ps.sign);
      }
    }
  }
  // Level-major work list.
  items_.clear();
  level_start_.clear();
  int maxnp = 0;
  for (const Tr& tr : tr_) maxnp = std::max(maxnp, tr.np);
  for (int lvl = 0; lvl < maxnp; ++lvl) {
    level_start_.push_back(items_.size());
    for (const Tr& tr : tr_)
      if (lvl < tr.np) {
        const uint32_t pi = (uint32_t)(tr.first + lvl);
        for (uint32_t k = 0; k < tiles[pi]; ++k) items_.push_back({pi, k});
      }
  }
  level_start_.push_back(items_.size());
  const int occ = std::max(1, sa_batch_occupancy(prec_, block_, lds_));
  grid_ = (int)std::min<int64_t>((int64_t)items_.size(),
                                 (int64_t)sched_.oversub * occ * cu_count());
  if (grid_ < 1) grid_ = 1;

  auto up = [](DevBuf& b, const void* h, size_t bytes) {
    if (b.bytes < bytes || !b.p) b = DevBuf(std::max<size_t>(bytes, 8));
    if (bytes) hip_check(hipMemcpy(b.p, h, bytes, hipMemcpyHostToDevice), "hipMemcpy(batch)");
  };
  for (int d = 0; d < 2; ++d) {
    up(d_pass_[d], desc[d].data(), np * sizeof(sa_pass_t));
    up(d_sign_[d], sign[d].data(), np);
  }
  up(d_tiles_, tiles.data(), np * sizeof(uint32_t));
  up(d_dep_, dep.data(), np * sizeof(int32_t));
  up(d_items_, items_.data(), items_.size() * sizeof(sa_batch_item_t));
  if (d_done_.bytes < std::max<size_t>(np, 1) * 4) d_done_ = DevBuf(std::max<size_t>(np, 1) * 4);
  if (!d_head_.p) {
    d_head_ = DevBuf(256);  // [0..7] queue heads, [8] wait-timeout flag (sticky)
    hip_check(hipMemset(d_head_.p, 0, 256), "hipMemset(batch flag)");
  }
  built_ = true;
}

void BatchPlan::copy_identities(hipStream_t s) {
  const int esz = prec_ ? 16 : 8;
  if (in_ == out_) return;
  for (const Tr& tr : tr_)
    if (tr.np == 0)
      hip_check(hipMemcpyAsync(static_cast<char*>(out_) + tr.off * esz,
                               static_cast<const char*>(in_) + tr.off * esz, tr.elems * esz,
                               hipMemcpyDeviceToDevice, s),
                "hipMemcpyAsync(batch identity)");
}

void BatchPlan::launch(int dir, hipStream_t s) {
  const int d = dir < 0 ? 0 : 1;
  copy_identities(s);
  if (items_.empty()) return;
  sa_batch_t b{};
  b.passes = static_cast<const sa_pass_t*>(d_pass_[d].p);
  b.pass_tiles = static_cast<const uint32_t*>(d_tiles_.p);
  b.pass_dep = static_cast<const int32_t*>(d_dep_.p);
  b.pass_sign = static_cast<const int8_t*>(d_sign_[d].p);
  b.done = static_cast<uint32_t*>(d_done_.p);
  b.head = static_cast<uint32_t*>(d_head_.p);
  b.err = static_cast<uint32_t*>(d_head_.p) + 8;
  // Two seconds of the device wall clock: far beyond any legitimate wait
  // for one pass of one transform.
  b.wait_ticks = 2ull * wall_clock_hz();
  b.mfma = mfma_;
  b.npasses = (uint32_t)prog_.npasses();
  const sa_batch_item_t* all = static_cast<const sa_batch_item_t*>(d_items_.p);
  if (!sched_.levels) {
    b.items = all;
    b.nitems = (uint32_t)items_.size();
    b.independent = 0;
    b.queues = 1;  // dependency waits need the list's order: one queue
    hip_check((hipError_t)sa_launch_batch(&b, prec_, grid_, block_, lds_, s), "batch launch");
    return;
  }
  // Level mode: one persistent launch per pass level, in stream order, one
  // work queue per XCD.
  b.independent = 1;
  for (size_t l = 0; l + 1 < level_start_.size(); ++l) {
    const size_t n = level_start_[l + 1] - level_start_[l];
    if (!n) continue;
    b.items = all + level_start_[l];
    b.nitems = (uint32_t)n;
    b.queues = batch_level_queues(n);
    const int g = (int)std::min<int64_t>((int64_t)n, grid_);
    hip_check((hipError_t)sa_launch_batch(&b, prec_, g, block_, lds_, s), "batch launch");
  }
}

void BatchPlan::set_precision_mode(int mode) {
  if (mode < 0 || mode > 2) fail(ERR_INVALID_DIM, "precision mode must be 0, 1 or 2");
  if (mode != 0 && prec_ != 0)
    fail(ERR_INVALID_FFTTYPE, "split-precision stages are for complex float (C2C) batches");
  if (mode == mfma_) return;
  mfma_ = mode;
  drop_graphs();  // captured launches carry the mode
}

void BatchPlan::execute(int dir, bool use_graph) {
  if (!built_) fail(ERR_NO_BUFFER, "batch buffers not set");
  if (!use_graph) {
    launch(dir, stream_);
    return;
  }
  const int d = dir < 0 ? 0 : 1;
  if (!graph_[d]) {
    hipStream_t cap = nullptr;
    hip_check(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    try {
      launch(dir, cap);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(cap, &g);
      if (g) (void)hipGraphDestroy(g);
      (void)hipStreamDestroy(cap);
      throw;
    }
    hip_check(hipStreamEndCapture(cap, &gdef_[d]), "hipStreamEndCapture");
    (void)hipStreamDestroy(cap);
    hip_check(hipGraphInstantiate(&graph_[d], gdef_[d], nullptr, nullptr, 0), "hipGraphInstantiate");
  }
  hip_check(hipGraphLaunch(graph_[d], stream_), "hipGraphLaunch");
}

// The wait-timeout flag is set by any launch since the last check (the
// kernels never clear it); reading it here clears it, so one timed-out
// launch fails exactly one synchronize.
void BatchPlan::check_timeout_flag() {
  if (!d_head_.p) return;
  uint32_t* flag = static_cast<uint32_t*>(d_head_.p) + 8;
  uint32_t err = 0;
  hip_check(hipMemcpy(&err, flag, 4, hipMemcpyDeviceToHost), "hipMemcpy(batch flag)");
  if (err) {
    hip_check(hipMemset(flag, 0, 4), "hipMemset(batch flag)");
    fail(ERR_BACKEND, "batch launch: a dependency wait timed out");
  }
}

void BatchPlan::synchronize() {
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  check_timeout_flag();
}

int64_t BatchPlan::total_bytes() const { return prog_.total_bytes(); }

std::string BatchPlan::autotune() {
  if (!built_) fail(ERR_NO_BUFFER, "batch buffers not set");
  const BatchSchedule cands[] = {{512, 1, true},  {256, 1, true}, {256, 2, true},
                                 {512, 1, false}, {256, 1, false}};
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0), "hipEventCreate");
  hip_check(hipEventCreate(&e1), "hipEventCreate");
  float best = 1e30f;
  BatchSchedule best_s = sched_;
  std::ostringstream os;
  for (const BatchSchedule& c : cands) {
    sched_ = c;
    drop_graphs();
    build();
    launch(-1, stream_);
    hip_check(hipEventRecord(e0, stream_), "hipEventRecord");
    for (int r = 0; r < 3; ++r) launch(-1, stream_);
    hip_check(hipEventRecord(e1, stream_), "hipEventRecord");
    hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    ms /= 3;
    check_timeout_flag();
    os << "block=" << c.block << " oversub=" << c.oversub << " mode=" << (c.levels ? "levels" : "deps")
       << " grid=" << grid_ << " : " << ms << " ms\n";
    if (ms < best) {
      best = ms;
      best_s = c;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  sched_ = best_s;
  drop_graphs();
  build();
  os << "chosen block=" << sched_.block << " oversub=" << sched_.oversub
     << " mode=" << (sched_.levels ? "levels" : "deps") << "\n";
  return os.str();
}

}  // namespace sa
