#include "server.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <sstream>

#include "batch.hpp"

namespace sa {

int64_t ServeJob::elems() const {
  if (kind == 1) return n;
  int64_t e = 1;
  for (int i = 0; i < ndim; ++i) e *= dims[i];
  return e;
}

std::string ServeJob::key() const {
  std::ostringstream os;
  os << (kind ? "1d" : "nd") << (prec ? "z" : "c");
  if (kind)
    os << ":" << n;
  else
    for (int i = 0; i < ndim; ++i) os << ":" << dims[i];
  return os.str();
}

// ---------------------------------------------------------------------------
Arena::Arena(int64_t bytes) : buf_((size_t)round(std::max<int64_t>(bytes, 0))) {
  size_ = (int64_t)buf_.bytes;
  if (size_ > 0) free_[0] = size_;
}

int64_t Arena::alloc(int64_t bytes) {
  const int64_t need = round(bytes);
  if (need == 0) return 0;
  for (auto it = free_.begin(); it != free_.end(); ++it) {
    if (it->second < need) continue;
    const int64_t off = it->first, rest = it->second - need;
    free_.erase(it);
    if (rest > 0) free_[off + need] = rest;
    used_ += need;
    peak_ = std::max(peak_, used_);
    return off;
  }
  return -1;
}

void Arena::release(int64_t off, int64_t bytes) {
  const int64_t len = round(bytes);
  if (len == 0) return;
  used_ -= len;
  auto it = free_.emplace(off, len).first;
  auto nx = std::next(it);
  if (nx != free_.end() && it->first + it->second == nx->first) {
    it->second += nx->second;
    free_.erase(nx);
  }
  if (it != free_.begin()) {
    auto pv = std::prev(it);
    if (pv->first + pv->second == it->first) {
      pv->second += it->second;
      free_.erase(it);
    }
  }
}

// ---------------------------------------------------------------------------
int64_t Server::now_ns() {
  return std::chrono::duration_cast<std::chrono::nanoseconds>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

Server::Server(const Config& cfg) : cfg_(cfg), arena_(cfg.arena_bytes) {
  if (cfg.slots < 1 || cfg.slots > 64) fail(ERR_INVALID_DIM, "slots must be in [1, 64]");
  if (cfg.planners < 1 || cfg.planners > 16) fail(ERR_INVALID_DIM, "planners must be in [1, 16]");
  if (cfg.max_bypass < 0) fail(ERR_INVALID_DIM, "max_bypass must be >= 0");
  if (cfg.mb_max_elems < 0 || cfg.mb_max_jobs < 1 || cfg.mb_max_inflight < 1)
    fail(ERR_INVALID_DIM, "invalid micro-batch limits");
  hip_check(hipGetDevice(&device_), "hipGetDevice");
  if (hipDeviceGetAttribute(&cus_, hipDeviceAttributeMultiprocessorCount, device_) != hipSuccess || cus_ < 1)
    cus_ = 256;
  slots_.resize(cfg.slots);
  try {
    for (Slot& s : slots_) {
      // Default priority; non-blocking so host-side table uploads on the
      // null stream never wait for served work.
      hip_check(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking), "hipStreamCreate");
      hip_check(hipEventCreate(&s.start), "hipEventCreate");
      hip_check(hipEventCreate(&s.end), "hipEventCreate");
      hip_check(hipHostMalloc(&s.mb_host, kMbStageBytes, hipHostMallocDefault), "hipHostMalloc");
      hip_check(hipMalloc(&s.mb_dev, kMbStageBytes), "hipMalloc");
    }
  } catch (...) {
    destroy_slots();
    throw;
  }
  for (int i = 0; i < cfg.planners; ++i) planners_.emplace_back([this] { planner(); });
  sched_ = std::thread([this] { scheduler(); });
}

void Server::destroy_slots() {
  for (Slot& s : slots_) {
    if (s.start) (void)hipEventDestroy(s.start);
    if (s.end) (void)hipEventDestroy(s.end);
    if (s.stream) (void)hipStreamDestroy(s.stream);
    if (s.mb_host) (void)hipHostFree(s.mb_host);
    if (s.mb_dev) (void)hipFree(s.mb_dev);
    s = Slot{};
  }
}

Server::~Server() {
  drain();
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  sched_cv_.notify_all();
  plan_cv_.notify_all();
  if (sched_.joinable()) sched_.join();
  for (auto& t : planners_)
    if (t.joinable()) t.join();
  for (Slot& s : slots_)
    if (s.stream) (void)hipStreamSynchronize(s.stream);
  destroy_slots();
}

int64_t Server::submit(const ServeJob& d, const void* in, void* out, int64_t arrival_ns) {
  if (d.kind != 0 && d.kind != 1) fail(ERR_INVALID_DIM, "job kind must be 0 (FFTND) or 1 (FFT1D)");
  if (d.prec != 0 && d.prec != 1) fail(ERR_INVALID_FFTTYPE, "invalid FFT type");
  if (d.dir != 0 && d.dir != 1) fail(ERR_INVALID_DIM, "direction must be 0 or 1");
  if (d.kind == 0) {
    if (d.ndim < 1 || d.ndim > 8) fail(ERR_INVALID_DIM, "ndim must be in [1, 8]");
    for (int i = 0; i < d.ndim; ++i)
      if (d.dims[i] <= 0) fail(ERR_INVALID_DIM, "dimension must be positive");
    int64_t e = 1;
    for (int i = 0; i < d.ndim; ++i) {
      if (e > INT64_MAX / 16 / d.dims[i]) fail(ERR_SIZE_OVERFLOW, "job size overflows");
      e *= d.dims[i];
    }
  } else if (d.n < 1 || d.n > INT64_MAX / 16) {
    fail(ERR_INVALID_DIM, "invalid FFT1D length");
  }
  if (!in || !out) fail(ERR_NULLPTR, "null job buffer");
  if (in == out) fail(ERR_INVALID_LAYOUT, "served jobs need distinct input and output buffers");
  const int64_t now = now_ns();
  std::lock_guard<std::mutex> lk(mu_);
  if (stop_) fail(ERR_INVALID_STATE, "server is shutting down");
  const int64_t id = next_id_++;
  Job& j = jobs_[id];
  j.id = id;
  j.d = d;
  j.in = in;
  j.out = out;
  j.t.arrival = arrival_ns < 0 ? now : arrival_ns;
  const std::string key = d.key();
  auto it = plans_.find(key);
  if (it == plans_.end()) {
    auto p = std::make_unique<Plan>();
    p->desc = d;
    j.plan = p.get();
    plans_.emplace(key, std::move(p));
    plan_queue_.push_back(key);
    plan_cv_.notify_one();
  } else {
    j.plan = it->second.get();
  }
  pending_.push_back(id);
  ++n_submitted_;
  dirty_ = true;
  sched_cv_.notify_one();
  return id;
}

void Server::planner() {
  (void)hipSetDevice(device_);
  std::unique_lock<std::mutex> lk(mu_);
  for (;;) {
    plan_cv_.wait(lk, [&] { return stop_ || !plan_queue_.empty(); });
    if (plan_queue_.empty()) return;  // stop_ and nothing left to build
    const std::string key = plan_queue_.front();
    plan_queue_.pop_front();
    Plan* p = plans_.at(key).get();
    const ServeJob d = p->desc;
    lk.unlock();
    std::unique_ptr<NDBackend> nd;
    std::unique_ptr<OneDBackend> od;
    int64_t scratch[2] = {0, 0};
    int status = 0;
    std::string msg;
    try {
      const int64_t esz = d.prec ? 16 : 8;
      if (d.kind == 0) {
        std::vector<int> ta(d.ndim);
        for (int i = 0; i < d.ndim; ++i) ta[i] = i;
        nd = std::make_unique<NDBackend>(d.ndim, ta.data(), d.ndim, d.dims, d.prec);
        for (int i = 0; i < 2; ++i) scratch[i] = nd->preserve_scratch_elems(i) * esz;
      } else {
        od = std::make_unique<OneDBackend>(d.n, d.prec, false);
        for (int i = 0; i < 2; ++i) scratch[i] = od->preserve_scratch_elems(i) * esz;
      }
    } catch (const Error& e) {
      status = e.status;
      msg = e.what();
    } catch (const std::exception& e) {
      status = ERR_INTERNAL;
      msg = e.what();
    }
    lk.lock();
    p->nd = std::move(nd);
    p->od = std::move(od);
    p->scratch[0] = scratch[0];
    p->scratch[1] = scratch[1];
    p->status = status;
    p->msg = msg;
    if (!status) build_mb(*p);
    p->state = status ? Plan::FAILED : Plan::READY;
    ++n_builds_;
    dirty_ = true;
    sched_cv_.notify_one();
  }
}

// Batch form of a plan: every pass of its input-preserving program re-tiled
// for the batch kernel (256-thread tiles; longer lines take up to 512).  A
// plan with a pass that has no batch form (the fixed-length Bluestein
// kernels, lines too long for one workgroup) keeps only the per-job path.
void Server::build_mb(Plan& p) const {
  const PassProgram& prog = p.nd ? p.nd->program() : p.od->program();
  p.prog = &prog;
  p.map = p.nd ? p.nd->preserve_map() : p.od->preserve_map();
  p.mb_ok = prog.npasses() > 0;
  const int esz = p.desc.prec ? 16 : 8;
  p.mb_pass.clear();
  p.mb_tiles.clear();
  p.mb_items = 0;
  p.mb_block = 64;
  p.mb_lds = 0;
  for (size_t i = 0; i < prog.npasses() && p.mb_ok; ++i) {
    const PassSpec& ps = prog.pass(i);
    sa_pass_t x = ps.p;
    if (ps.blue || !batch_retile(x, esz, 256, kBatchLdsBudget)) {
      p.mb_ok = false;
      break;
    }
    const uint32_t tiles = (uint32_t)((x.batch + x.C - 1) / x.C);
    p.mb_pass.push_back(x);
    p.mb_tiles.push_back(tiles);
    p.mb_items += tiles;
    p.mb_block = std::max(p.mb_block, (x.C * x.T + 63) / 64 * 64);
    p.mb_lds = std::max(p.mb_lds, (size_t)x.C * x.ldsstride * esz);
  }
  if (p.mb_ok && mb_stage_bytes((int64_t)p.mb_pass.size(), p.mb_items) > kMbStageBytes) p.mb_ok = false;
}

static size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Staging layout of one micro-batch: descriptors, signs, tiles per pass,
// the pass each pass follows, per-pass counters, eight queue heads per pass
// level (zero, so no clearing launch runs) + the timeout flag, work items.
size_t Server::mb_stage_bytes(int64_t np, int64_t items) {
  return align256((size_t)np * sizeof(sa_pass_t)) + align256((size_t)np) + 3 * align256((size_t)np * 4) +
         align256((size_t)(8 * np + 1) * 4) + (size_t)items * sizeof(sa_batch_item_t);
}

bool Server::mb_eligible(const Job& j, const Plan& p) const {
  return p.mb_ok && !j.d.normalize && j.d.elems() <= cfg_.mb_max_elems;
}

void Server::set_batching(int64_t max_elems, int max_jobs, int max_inflight) {
  if (max_elems < 0 || max_jobs < 1 || max_inflight < 1) fail(ERR_INVALID_DIM, "invalid micro-batch limits");
  std::lock_guard<std::mutex> lk(mu_);
  cfg_.mb_max_elems = max_elems;
  cfg_.mb_max_jobs = max_jobs;
  cfg_.mb_max_inflight = max_inflight;
}

void Server::release(Job& j) {
  for (int i = 0; i < 2; ++i)
    if (j.off[i] >= 0) {
      arena_.release(j.off[i], j.bytes[i]);
      j.off[i] = -1;
    }
  if (j.slot >= 0) {
    Slot& s = slots_[j.slot];
    if (s.job == kSlotBatch) {
      s.mb_jobs.erase(std::remove(s.mb_jobs.begin(), s.mb_jobs.end(), j.id), s.mb_jobs.end());
      if (s.mb_jobs.empty()) s.job = kSlotFree;
    } else {
      s.job = kSlotFree;
    }
    j.slot = -1;
    --running_;
  }
}

void Server::finish(Job& j, int status, const std::string& msg) {
  if (j.state == RUNNING) release(j);
  j.state = DONE;
  j.status = status;
  j.msg = msg;
  if (!j.t.done) j.t.done = now_ns();
  if (status == 0) {
    ++n_completed_;
    lat_ns_.push_back(j.t.done - j.t.arrival);
    first_arrival_ = std::min(first_arrival_, j.t.arrival);
    last_done_ = std::max(last_done_, j.t.done);
    const double n = (double)j.d.elems();
    if (n > 1) gflop_ += 5.0 * n * std::log2(n) * 1e-9;
  } else {
    ++n_failed_;
  }
  dirty_ = true;
  done_cv_.notify_all();
}

// Launch one admitted job on its slot (scheduler thread, lock released).
// Returns 0 or the failure status (message in msg).
int Server::launch(Job& j, const Plan& p, Slot& s, std::string& msg) {
  try {
    void* s1 = j.off[0] >= 0 ? arena_.base() + j.off[0] : nullptr;
    void* s2 = j.off[1] >= 0 ? arena_.base() + j.off[1] : nullptr;
    const int dir = j.d.dir == 0 ? -1 : +1;
    hip_check(hipEventRecord(s.start, s.stream), "hipEventRecord");
    if (p.nd)
      p.nd->execute_preserve(j.in, j.out, s1, s2, dir, s.stream);
    else
      p.od->execute_preserve(j.in, j.out, s1, s2, dir, s.stream);
    if (j.d.normalize) {
      const long double f = 1.0L / std::sqrt((long double)j.d.elems());
      scale_async(j.out, (double)f, (uint64_t)j.d.elems(), j.d.prec, s.stream);
    }
    hip_check(hipEventRecord(s.end, s.stream), "hipEventRecord");
    return 0;
  } catch (const Error& e) {
    msg = e.what();
    (void)hipStreamSynchronize(s.stream);  // nothing may still use the scratch
    return e.status;
  } catch (const std::exception& e) {
    msg = e.what();
    (void)hipStreamSynchronize(s.stream);
    return ERR_INTERNAL;
  }
}

// Launch one micro-batch on its slot (scheduler thread, lock released):
// descriptors with each job's buffers and direction, the level-major work
// list, one copy to the device, then one batch-kernel launch per pass level
// (independent mode: the stream orders the levels).  Returns 0 or the
// failure status of the whole group.
int Server::launch_mb(Slot& s, const std::vector<Job*>& js, std::string& msg) {
  try {
    const int prec = js[0]->d.prec;
    int64_t np = 0, items = 0;
    int block = 64, maxnp = 0;
    size_t lds = 0;
    for (const Job* j : js) {
      const Plan& p = *j->plan;
      np += (int64_t)p.mb_pass.size();
      items += p.mb_items;
      block = std::max(block, p.mb_block);
      lds = std::max(lds, p.mb_lds);
      maxnp = std::max(maxnp, (int)p.mb_pass.size());
    }
    const size_t bytes = mb_stage_bytes(np, items);
    if (bytes > kMbStageBytes) fail(ERR_INTERNAL, "micro-batch exceeds its staging");
    char* h = static_cast<char*>(s.mb_host);
    const size_t o_sign = align256((size_t)np * sizeof(sa_pass_t));
    const size_t o_tiles = o_sign + align256((size_t)np);
    const size_t o_dep = o_tiles + align256((size_t)np * 4);
    const size_t o_done = o_dep + align256((size_t)np * 4);
    const size_t o_head = o_done + align256((size_t)np * 4);  // [8 heads per level..., timeout flag]
    const size_t o_items = o_head + align256((size_t)(8 * np + 1) * 4);
    sa_pass_t* desc = reinterpret_cast<sa_pass_t*>(h);
    int8_t* sign = reinterpret_cast<int8_t*>(h + o_sign);
    uint32_t* tiles = reinterpret_cast<uint32_t*>(h + o_tiles);
    int32_t* dep = reinterpret_cast<int32_t*>(h + o_dep);
    std::memset(h + o_done, 0, o_items - o_done);  // counters, head, timeout flag
    std::vector<uint32_t> first(js.size());
    uint32_t k = 0;
    for (size_t q = 0; q < js.size(); ++q) {
      const Job& j = *js[q];
      const Plan& p = *j.plan;
      void* s1 = j.off[0] >= 0 ? arena_.base() + j.off[0] : nullptr;
      void* s2 = j.off[1] >= 0 ? arena_.base() + j.off[1] : nullptr;
      void* bufs[SLOT_COUNT] = {const_cast<void*>(j.in), j.out, s1, s2};
      const int dir = j.d.dir == 0 ? -1 : +1;
      const int d = dir < 0 ? 0 : 1;
      first[q] = k;
      for (size_t i = 0; i < p.mb_pass.size(); ++i, ++k) {
        sa_pass_t x = p.mb_pass[i];
        x.in = bufs[p.map.in[i]];
        x.out = bufs[p.map.out[i]];
        if (!x.in || !x.out) fail(ERR_NO_BUFFER, "micro-batch pass buffer not set");
        const PassSpec& ps = p.prog->pass(i);
        if (ps.dir_tables) {
          x.pre = ps.pre_t[d];
          x.post = ps.post_t[d];
          x.blu_h = ps.blu_t[d];
        }
        desc[k] = x;
        sign[k] = (int8_t)(ps.sign == 0 ? dir : ps.sign);
        tiles[k] = p.mb_tiles[i];
        dep[k] = i ? (int32_t)k - 1 : -1;  // ordered by the level launches, not waited on
      }
    }
    sa_batch_item_t* it = reinterpret_cast<sa_batch_item_t*>(h + o_items);
    std::vector<size_t> level(maxnp + 1);
    size_t ni = 0;
    for (int l = 0; l < maxnp; ++l) {
      level[l] = ni;
      for (size_t q = 0; q < js.size(); ++q) {
        if (l >= (int)js[q]->plan->mb_pass.size()) continue;
        const uint32_t pi = first[q] + (uint32_t)l;
        for (uint32_t t = 0; t < tiles[pi]; ++t) it[ni++] = {pi, t};
      }
    }
    level[maxnp] = ni;
    char* dv = static_cast<char*>(s.mb_dev);
    hip_check(hipEventRecord(s.start, s.stream), "hipEventRecord");
    hip_check(hipMemcpyAsync(dv, h, o_items + ni * sizeof(sa_batch_item_t), hipMemcpyHostToDevice,
                             s.stream),
              "hipMemcpyAsync(micro-batch)");
    const auto key = std::make_pair(prec, std::make_pair(block, lds));
    auto oc = occ_cache_.find(key);
    if (oc == occ_cache_.end())
      oc = occ_cache_.emplace(key, std::max(1, sa_batch_occupancy(prec, block, lds))).first;
    const int64_t cap = (int64_t)oc->second * cus_;
    sa_batch_t b{};
    b.passes = reinterpret_cast<const sa_pass_t*>(dv);
    b.pass_sign = reinterpret_cast<const int8_t*>(dv + o_sign);
    b.pass_tiles = reinterpret_cast<const uint32_t*>(dv + o_tiles);
    b.pass_dep = reinterpret_cast<const int32_t*>(dv + o_dep);
    b.wait_ticks = 1;  // independent levels never wait
    b.done = reinterpret_cast<uint32_t*>(dv + o_done);
    uint32_t* heads = reinterpret_cast<uint32_t*>(dv + o_head);
    b.err = heads + 8 * np;  // levels <= passes
    b.npasses = (uint32_t)np;
    b.independent = 1;
    b.mfma = 0;
    b.precleared = 1;  // every level has its own zeroed heads; `done` is unused when independent

    const sa_batch_item_t* dit = reinterpret_cast<const sa_batch_item_t*>(dv + o_items);
    for (int l = 0; l < maxnp; ++l) {
      const size_t n = level[l + 1] - level[l];
      if (!n) continue;
      b.items = dit + level[l];
      b.nitems = (uint32_t)n;
      b.queues = batch_level_queues(n);  // per-XCD queues on long levels
      b.head = heads + 8 * l;
      const int g = (int)std::min<int64_t>((int64_t)n, cap);
      hip_check((hipError_t)sa_launch_batch(&b, prec, g, block, lds, s.stream), "micro-batch launch");
    }
    hip_check(hipEventRecord(s.end, s.stream), "hipEventRecord");
    return 0;
  } catch (const Error& e) {
    msg = e.what();
    (void)hipStreamSynchronize(s.stream);
    return e.status;
  } catch (const std::exception& e) {
    msg = e.what();
    (void)hipStreamSynchronize(s.stream);
    return ERR_INTERNAL;
  }
}

void Server::scheduler() {
  (void)hipSetDevice(device_);
  std::unique_lock<std::mutex> lk(mu_);
  std::vector<int64_t> admitted;
  std::vector<int> admitted_mb;  // slots of the micro-batches formed in this loop iteration
  for (;;) {
    bool progress = false;
    // Retire: a slot's job is done when its end event has completed.
    for (Slot& s : slots_) {
      if (s.job == kSlotFree || s.launching) continue;
      const hipError_t e = hipEventQuery(s.end);
      if (e == hipErrorNotReady) continue;
      const int64_t t_done = now_ns();
      int64_t gpu_ns = 0;
      if (e == hipSuccess) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, s.start, s.end) == hipSuccess) gpu_ns = (int64_t)((double)ms * 1e6);
      }
      // A micro-batch's jobs all end with its last launch.
      const std::vector<int64_t> ids = s.job == kSlotBatch ? s.mb_jobs : std::vector<int64_t>{s.job};
      for (const int64_t id : ids) {
        Job& j = jobs_.at(id);
        j.t.done = t_done;
        j.t.gpu_ns = gpu_ns;
        if (e == hipSuccess)
          finish(j, 0, std::string());
        else
          finish(j, ERR_BACKEND, std::string("served job: ") + hipGetErrorString(e));
      }
      progress = true;
    }
    // Admit in arrival order while slots and arena space last.
    admitted.clear();
    admitted_mb.clear();
    if (dirty_) {
      dirty_ = false;
      std::vector<Job*> blocked;
      int free_slots = 0;
      int mb_running[2] = {0, 0};  // micro-batches on the device per precision
      for (const Slot& s : slots_) {
        free_slots += s.job == kSlotFree;
        if (s.job == kSlotBatch) ++mb_running[s.mb_prec];
      }
      int open_mb[2] = {-1, -1};  // micro-batch being formed per precision (slot)
      for (auto it = pending_.begin(); it != pending_.end();) {
        if (free_slots == 0 && open_mb[0] < 0 && open_mb[1] < 0) break;
        Job& j = jobs_.at(*it);
        const Plan& p = *j.plan;
        if (p.state == Plan::BUILDING) {
          ++it;
          continue;
        }
        if (p.state == Plan::FAILED) {
          finish(j, p.status, p.msg);
          it = pending_.erase(it);
          continue;
        }
        if (Arena::round(p.scratch[0]) + Arena::round(p.scratch[1]) > arena_.size()) {
          finish(j, ERR_ALLOC, "job scratch (" + std::to_string(p.scratch[0] + p.scratch[1]) +
                                   " bytes) exceeds the server arena");
          it = pending_.erase(it);
          continue;
        }
        bool hol = false;  // an earlier job has been passed too often
        for (const Job* b : blocked) hol |= b->bypassed >= cfg_.max_bypass;
        if (hol) break;
        const bool mb = mb_eligible(j, p);
        const int pr = j.d.prec;
        if (mb && open_mb[pr] >= 0) {
          const Slot& o = slots_[open_mb[pr]];
          if ((int)o.mb_jobs.size() >= cfg_.mb_max_jobs ||
              mb_stage_bytes(o.mb_np + (int64_t)p.mb_pass.size(), o.mb_items + p.mb_items) > kMbStageBytes)
            open_mb[pr] = -1;  // full: it launches as it is
        }
        if (free_slots == 0 && !(mb && open_mb[pr] >= 0)) {
          ++it;  // needs a slot of its own
          continue;
        }
        if (mb && open_mb[pr] < 0 && mb_running[pr] >= cfg_.mb_max_inflight) {
          ++it;  // waits for a running group of its precision: the next group gathers more
          continue;
        }
        int64_t off0 = arena_.alloc(p.scratch[0]), off1 = -1;
        if (off0 >= 0) {
          off1 = arena_.alloc(p.scratch[1]);
          if (off1 < 0) {
            arena_.release(off0, p.scratch[0]);
            off0 = -1;
          }
        }
        if (off0 < 0) {
          blocked.push_back(&j);
          ++it;
          continue;
        }
        for (Job* b : blocked) ++b->bypassed;
        int si = 0;
        if (mb && open_mb[pr] >= 0) {
          si = open_mb[pr];
        } else {
          while (slots_[si].job != kSlotFree) ++si;
          slots_[si].launching = true;
          --free_slots;
          if (mb) {
            Slot& o = slots_[si];
            o.job = kSlotBatch;
            o.mb_jobs.clear();
            o.mb_prec = pr;
            o.mb_np = o.mb_items = 0;
            open_mb[pr] = si;
            admitted_mb.push_back(si);
            ++mb_running[pr];
            ++n_batches_;
          } else {
            slots_[si].job = *it;
            admitted.push_back(*it);
          }
        }
        if (mb) {
          Slot& o = slots_[si];
          o.mb_jobs.push_back(*it);
          o.mb_np += (int64_t)p.mb_pass.size();
          o.mb_items += p.mb_items;
          ++n_batched_jobs_;
        }
        j.slot = si;
        j.off[0] = p.scratch[0] ? off0 : -1;
        j.off[1] = p.scratch[1] ? off1 : -1;
        j.bytes[0] = p.scratch[0];
        j.bytes[1] = p.scratch[1];
        j.state = RUNNING;
        j.t.admit = now_ns();
        ++running_;
        it = pending_.erase(it);
      }
      progress = true;
    }
    if (!admitted.empty() || !admitted_mb.empty()) {
      // Job and plan objects are node-stable and only this thread retires
      // running jobs, so the references stay valid without the lock.
      std::vector<std::pair<int, std::string>> res(admitted.size()), res_mb(admitted_mb.size());
      std::vector<Job*> js(admitted.size());
      std::vector<std::vector<Job*>> groups(admitted_mb.size());
      for (size_t k = 0; k < admitted.size(); ++k) js[k] = &jobs_.at(admitted[k]);
      for (size_t k = 0; k < admitted_mb.size(); ++k)
        for (const int64_t id : slots_[admitted_mb[k]].mb_jobs) groups[k].push_back(&jobs_.at(id));
      lk.unlock();
      for (size_t k = 0; k < admitted_mb.size(); ++k)
        res_mb[k].first = launch_mb(slots_[admitted_mb[k]], groups[k], res_mb[k].second);
      for (size_t k = 0; k < admitted.size(); ++k)
        res[k].first = launch(*js[k], *js[k]->plan, slots_[js[k]->slot], res[k].second);
      lk.lock();
      for (size_t k = 0; k < admitted.size(); ++k) {
        slots_[js[k]->slot].launching = false;
        if (res[k].first) finish(*js[k], res[k].first, res[k].second);
      }
      for (size_t k = 0; k < admitted_mb.size(); ++k) {
        slots_[admitted_mb[k]].launching = false;
        if (res_mb[k].first)
          for (Job* g : groups[k]) finish(*g, res_mb[k].first, res_mb[k].second);
      }
    }
    if (stop_ && running_ == 0 && pending_.empty()) break;
    if (!progress) {
      if (running_ > 0) {
        lk.unlock();
        std::this_thread::yield();
        lk.lock();
      } else {
        sched_cv_.wait(lk, [&] { return stop_ || dirty_; });
      }
    }
  }
}

int Server::wait(int64_t id, int64_t timeout_ns, int* status, ServeTimes* t, std::string* msg) {
  std::unique_lock<std::mutex> lk(mu_);
  auto it = jobs_.find(id);
  if (it == jobs_.end()) fail(ERR_INVALID_STATE, "unknown job id " + std::to_string(id));
  Job& j = it->second;
  auto ready = [&] { return j.state == DONE; };
  if (timeout_ns < 0) {
    done_cv_.wait(lk, ready);
  } else if (!done_cv_.wait_for(lk, std::chrono::nanoseconds(timeout_ns), ready)) {
    return 1;
  }
  if (status) *status = j.status;
  if (t) *t = j.t;
  if (msg) *msg = j.msg;
  jobs_.erase(it);
  return 0;
}

int Server::query(int64_t id) const {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = jobs_.find(id);
  if (it == jobs_.end()) fail(ERR_INVALID_STATE, "unknown job id " + std::to_string(id));
  return it->second.state == DONE ? 0 : 1;
}

void Server::drain() {
  std::unique_lock<std::mutex> lk(mu_);
  done_cv_.wait(lk, [&] { return pending_.empty() && running_ == 0; });
}

ServeStats Server::stats() const {
  std::lock_guard<std::mutex> lk(mu_);
  ServeStats s;
  s.submitted = n_submitted_;
  s.completed = n_completed_;
  s.failed = n_failed_;
  s.plans = (int64_t)plans_.size();
  s.plan_builds = n_builds_;
  s.arena_bytes = arena_.size();
  s.arena_peak = arena_.peak();
  for (const auto& kv : plans_) {
    const Plan& p = *kv.second;
    if (p.state != Plan::READY) continue;
    s.table_bytes += p.nd ? p.nd->table_bytes() : p.od->table_bytes();
  }
  if (!lat_ns_.empty()) {
    std::vector<int64_t> v(lat_ns_);
    std::sort(v.begin(), v.end());
    auto rank = [&](double q) {  // nearest-rank percentile
      size_t k = (size_t)std::ceil(q * (double)v.size());
      return (double)v[k ? k - 1 : 0] * 1e-6;
    };
    s.p50_ms = rank(0.50);
    s.p99_ms = rank(0.99);
    double sum = 0;
    for (int64_t x : v) sum += (double)x;
    s.mean_ms = sum / (double)v.size() * 1e-6;
    s.max_ms = (double)v.back() * 1e-6;
    s.span_s = (double)(last_done_ - first_arrival_) * 1e-9;
  }
  s.gflop = gflop_;
  s.batches = n_batches_;
  s.batched_jobs = n_batched_jobs_;
  return s;
}

void Server::reset_stats() {
  std::lock_guard<std::mutex> lk(mu_);
  n_submitted_ = n_completed_ = n_failed_ = 0;
  n_batches_ = n_batched_jobs_ = 0;
  lat_ns_.clear();
  first_arrival_ = INT64_MAX;
  last_done_ = 0;
  gflop_ = 0;
  arena_.reset_peak();
}

}  // namespace sa
