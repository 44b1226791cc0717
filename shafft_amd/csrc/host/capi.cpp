// C ABI: backend seam (shafftAmd*) and the reference's public C API at one
// rank per GPU (shafft*).  The plan state machine, normalisation exponent
// bookkeeping and buffer-swap contract follow the reference core:
//   state machine      src/cpp/shafft.cpp:55-75,172-186,913-983
//   configure checks   src/cpp/shafft_internal.cpp:229-333
//   execute/normalize  src/cpp/shafft_internal.cpp:62-188 (ND),
//                      src/cpp/shafft_internal.cpp:584-643 (1D)
//   1-D factor         src/cpp/fft/hipfft_method/fft1d_method.hip:652-679
#include <hip/hip_runtime_api.h>
#include <unistd.h>

#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <random>
#include <vector>

#include "../../../include/shafft_amd.h"
#include "backend.hpp"
#include "batch.hpp"
#include "common.hpp"
#include "server.hpp"

using namespace sa;

#define SHAFFT_AMD_VERSION_MAJOR 1
#define SHAFFT_AMD_VERSION_MINOR 1
#define SHAFFT_AMD_VERSION_PATCH 0

struct shafft_amd_nd_plan {
  std::unique_ptr<NDBackend> be;
};
struct shafft_amd_1d_plan {
  std::unique_ptr<OneDBackend> be;
};
struct shafft_amd_batch_plan {
  std::unique_ptr<BatchPlan> be;
};
struct shafft_amd_server {
  std::unique_ptr<Server> be;
};

namespace {

constexpr uint32_t kMagic = 0x5AFF7A3Du;

enum class Kind { ND, OneD };
enum class State { UNINITIALIZED, CONFIGURED, PLANNED };

// Plan object behind the reference C API's void* handles.
struct Plan {
  uint32_t magic = kMagic;
  Kind kind;
  State state = State::UNINITIALIZED;
  int prec = 0;
  std::vector<int> dims;  // ND extents (1D: {N})
  size_t global = 0;
  long double norm_denominator = 1.0L;
  int norm_exponent = 0;
  void* data = nullptr;
  void* work = nullptr;
  hipStream_t stream = nullptr;
  int output_policy = 0;
  int layout_state = 0;  // 1D: 0 INITIAL, 1 REDISTRIBUTED
  std::unique_ptr<NDBackend> nd;
  std::unique_ptr<OneDBackend> od;
  explicit Plan(Kind k) : kind(k) {}
};

Plan* as_plan(void* p) {
  if (!p) fail(ERR_NULLPTR, "null plan");
  Plan* q = static_cast<Plan*>(p);
  if (q->magic != kMagic) fail(ERR_NULLPTR, "not a plan handle");
  return q;
}

void check_comm(shafft_comm_t comm) {
  if (comm != 0) fail(ERR_INVALID_COMM, "only the one-rank communicator (0) is supported");
}

int prec_of(shafft_t t) {
  if (t == SHAFFT_C2C) return 0;
  if (t == SHAFFT_Z2Z) return 1;
  fail(ERR_INVALID_FFTTYPE, "invalid FFT type");
}

}  // namespace

// =========================================================================
// 1. Backend seam
// =========================================================================
extern "C" {

int shafftAmdNDPlan(shafft_amd_nd_plan** plan, int nta, const int* ta, int ndim,
                    const int* dimensions, int precision, void* in, void* out) {
  (void)in;
  (void)out;
  return guarded([&] {
    if (!plan) fail(ERR_NULLPTR, "null plan pointer");
    *plan = nullptr;
    auto p = std::make_unique<shafft_amd_nd_plan>();
    p->be = std::make_unique<NDBackend>(nta, ta, ndim, dimensions, precision);
    *plan = p.release();
    return 0;
  });
}

int shafftAmdNDSetStream(shafft_amd_nd_plan* plan, void* stream) {
  return guarded([&] {
    if (!plan) fail(ERR_NULLPTR, "null plan");
    plan->be->set_stream(reinterpret_cast<hipStream_t>(stream));
    return 0;
  });
}

int shafftAmdNDExecute(shafft_amd_nd_plan* plan, void** data, void** work, int direction) {
  return guarded([&] {
    if (!plan || !data || !work) fail(ERR_NULLPTR, "null argument");
    plan->be->execute(*data, *work, direction == 0 ? -1 : +1);
    return 0;
  });
}

int shafftAmdNDDestroy(shafft_amd_nd_plan* plan) {
  return guarded([&] {
    delete plan;
    return 0;
  });
}

int shafftAmdNDFinalize(void) { return 0; }

int shafftAmdNormalizeComplexFloat(float factor, size_t tensor_size, void* data, void* stream) {
  return guarded([&] {
    scale_async(data, factor, tensor_size, 0, reinterpret_cast<hipStream_t>(stream));
    return 0;
  });
}

int shafftAmdNormalizeComplexDouble(double factor, size_t tensor_size, void* data, void* stream) {
  return guarded([&] {
    scale_async(data, factor, tensor_size, 1, reinterpret_cast<hipStream_t>(stream));
    return 0;
  });
}

int shafftAmd1DQueryLayout(size_t globalN, size_t* localN, size_t* localStart,
                           size_t* localAllocSize, size_t* localNTrans,
                           size_t* localStartTrans) {
  return guarded([&] {
    if (!localN || !localStart || !localAllocSize) fail(ERR_NULLPTR, "null output");
    if (globalN == 0) fail(ERR_INVALID_DIM, "N must be positive");
    // One rank: L = N, the whole sequence is local before and after.
    *localN = globalN;
    *localStart = 0;
    *localAllocSize = globalN;
    if (localNTrans) *localNTrans = globalN;
    if (localStartTrans) *localStartTrans = 0;
    return 0;
  });
}

int shafftAmd1DPlan(shafft_amd_1d_plan** plan, size_t globalN, int precision, void* in,
                    void* out) {
  (void)in;
  (void)out;
  return guarded([&] {
    if (!plan) fail(ERR_NULLPTR, "null plan pointer");
    *plan = nullptr;
    if (globalN == 0) fail(ERR_INVALID_DIM, "N must be positive");
    auto p = std::make_unique<shafft_amd_1d_plan>();
    p->be = std::make_unique<OneDBackend>((int64_t)globalN, precision);
    *plan = p.release();
    return 0;
  });
}

int shafftAmd1DExecute(shafft_amd_1d_plan* plan, void* in, void* out, int direction) {
  return guarded([&] {
    if (!plan) fail(ERR_NULLPTR, "null plan");
    plan->be->execute(in, out, direction == 0 ? -1 : +1);
    return 0;
  });
}

int shafftAmd1DExecuteNoAlloc(shafft_amd_1d_plan* plan, const void* in, void* out,
                              int direction) {
  return guarded([&] {
    if (!plan) fail(ERR_NULLPTR, "null plan");
    void* scratch = nullptr;
    size_t bytes = 0;
    hip_check((hipError_t)sa_noalloc_scratch(&scratch, &bytes), "allocation-free scratch");
    plan->be->execute_in_scratch(in, out, direction == 0 ? -1 : +1, scratch, bytes);
    return 0;
  });
}

int shafftAmd1DSetStream(shafft_amd_1d_plan* plan, void* stream) {
  return guarded([&] {
    if (!plan) fail(ERR_NULLPTR, "null plan");
    plan->be->set_stream(reinterpret_cast<hipStream_t>(stream));
    return 0;
  });
}

int shafftAmd1DSynchronize(shafft_amd_1d_plan* plan) {
  return guarded([&] {
    if (!plan) fail(ERR_NULLPTR, "null plan");
    hip_check(hipStreamSynchronize(plan->be->stream()), "hipStreamSynchronize");
    return 0;
  });
}

int shafftAmd1DDestroy(shafft_amd_1d_plan* plan) {
  return guarded([&] {
    delete plan;
    return 0;
  });
}

int shafftAmd1DNormalize(shafft_amd_1d_plan* plan, void* data, int normExponent,
                         size_t localCount) {
  return guarded([&] {
    if (!plan) fail(ERR_PLAN_NOT_INIT, "null plan");
    const size_t n = (size_t)plan->be->n();
    double f;
    if (plan->be->prec() == 0) {
      // Working precision, as fft1d_method.hip:663-668.
      float norm = 1.0f / std::pow(std::sqrt(static_cast<float>(n)), normExponent);
      f = norm;
    } else {
      f = 1.0 / std::pow(std::sqrt(static_cast<double>(n)), normExponent);
    }
    scale_async(data, f, localCount, plan->be->prec(), plan->be->stream());
    hip_check(hipStreamSynchronize(plan->be->stream()), "hipStreamSynchronize");
    return 0;
  });
}

int shafftAmdNDNumPasses(shafft_amd_nd_plan* plan, int* npasses) {
  return guarded([&] {
    if (!plan || !npasses) fail(ERR_NULLPTR, "null argument");
    *npasses = (int)plan->be->program().npasses();
    return 0;
  });
}

int shafftAmdNDPassBytes(shafft_amd_nd_plan* plan, int pass, int64_t* bytes) {
  return guarded([&] {
    if (!plan || !bytes) fail(ERR_NULLPTR, "null argument");
    if (pass < 0 || (size_t)pass >= plan->be->program().npasses())
      fail(ERR_INVALID_DIM, "pass index out of range");
    *bytes = plan->be->program().pass((size_t)pass).bytes;
    return 0;
  });
}

int shafftAmdNDExecutePass(shafft_amd_nd_plan* plan, int pass, void* in, int64_t in_elems,
                           void* out, int64_t out_elems, int direction) {
  return guarded([&] {
    if (!plan) fail(ERR_NULLPTR, "null plan");
    if (pass < 0) fail(ERR_INVALID_DIM, "pass index out of range");
    plan->be->execute_pass((size_t)pass, in, in_elems, out, out_elems, direction == 0 ? -1 : +1);
    return 0;
  });
}

int shafftAmdNDPassExtents(shafft_amd_nd_plan* plan, int pass, int64_t* in_elems,
                           int64_t* out_elems) {
  return guarded([&] {
    if (!plan || !in_elems || !out_elems) fail(ERR_NULLPTR, "null argument");
    if (pass < 0) fail(ERR_INVALID_DIM, "pass index out of range");
    plan->be->pass_extents((size_t)pass, in_elems, out_elems);
    return 0;
  });
}

int shafftAmd1DExecutePass(shafft_amd_1d_plan* plan, int pass, void* in, int64_t in_elems,
                           void* out, int64_t out_elems, int direction) {
  return guarded([&] {
    if (!plan) fail(ERR_NULLPTR, "null plan");
    if (pass < 0) fail(ERR_INVALID_DIM, "pass index out of range");
    plan->be->execute_pass((size_t)pass, in, in_elems, out, out_elems, direction == 0 ? -1 : +1);
    return 0;
  });
}

int shafftAmd1DPassExtents(shafft_amd_1d_plan* plan, int pass, int64_t* in_elems,
                           int64_t* out_elems) {
  return guarded([&] {
    if (!plan || !in_elems || !out_elems) fail(ERR_NULLPTR, "null argument");
    if (pass < 0) fail(ERR_INVALID_DIM, "pass index out of range");
    plan->be->pass_extents((size_t)pass, in_elems, out_elems);
    return 0;
  });
}

int shafftAmd1DNumPasses(shafft_amd_1d_plan* plan, int* npasses) {
  return guarded([&] {
    if (!plan || !npasses) fail(ERR_NULLPTR, "null argument");
    *npasses = (int)plan->be->program().npasses();
    return 0;
  });
}

int shafftAmd1DPassBytes(shafft_amd_1d_plan* plan, int pass, int64_t* bytes) {
  return guarded([&] {
    if (!plan || !bytes) fail(ERR_NULLPTR, "null argument");
    if (pass < 0 || (size_t)pass >= plan->be->program().npasses())
      fail(ERR_INVALID_DIM, "pass index out of range");
    *bytes = plan->be->program().pass((size_t)pass).bytes;
    return 0;
  });
}

static int copy_desc(const std::string& s, char* buf, int buflen) {
  if (!buf || buflen <= 0) fail(ERR_NULLPTR, "null buffer");
  const size_t n = std::min<size_t>(s.size(), (size_t)buflen - 1);
  std::memcpy(buf, s.data(), n);
  buf[n] = 0;
  return 0;
}

int shafftAmdNDDescribe(shafft_amd_nd_plan* plan, char* buf, int buflen) {
  return guarded([&] {
    if (!plan) fail(ERR_NULLPTR, "null plan");
    return copy_desc(plan->be->program().describe(), buf, buflen);
  });
}

int shafftAmd1DDescribe(shafft_amd_1d_plan* plan, char* buf, int buflen) {
  return guarded([&] {
    if (!plan) fail(ERR_NULLPTR, "null plan");
    return copy_desc(plan->be->program().describe(), buf, buflen);
  });
}

// Synthetic inputs of the benchmark workloads: std::mt19937_64(seed), the
// generator the reference's tests draw from (tests/integration/cpp/
// test_correctness.cpp:382), with the value transform spelled out rather
// than left to std::*_distribution (implementation-defined):
//   distribution 0: per element re then im, u = (x >> 11) 2^-53, value 2u - 1;
//   distribution 1: Box-Muller per element, u1 = ((x0 >> 11) + 1) 2^-53,
//                   u2 = (x1 >> 11) 2^-53, r = sqrt(-2 ln u1),
//                   value r (cos 2 pi u2 + i sin 2 pi u2).
// Values are formed in double and rounded once for complex float.
int shafftAmdSyntheticInput(void* host_out, int64_t count, uint64_t seed, int distribution,
                            int precision) {
  return guarded([&] {
    if (!host_out && count > 0) fail(ERR_NULLPTR, "null output");
    if (count < 0) fail(ERR_INVALID_DIM, "negative count");
    if (distribution != 0 && distribution != 1) fail(ERR_INVALID_DIM, "unknown distribution");
    if (precision != 0 && precision != 1) fail(ERR_INVALID_FFTTYPE, "invalid FFT type");
    std::mt19937_64 g(seed);
    const double k53 = 1.0 / 9007199254740992.0, two_pi = 6.283185307179586476925286766559;
    for (int64_t i = 0; i < count; ++i) {
      double re, im;
      if (distribution == 0) {
        re = 2.0 * ((double)(g() >> 11) * k53) - 1.0;
        im = 2.0 * ((double)(g() >> 11) * k53) - 1.0;
      } else {
        const double u1 = (double)((g() >> 11) + 1) * k53;
        const double u2 = (double)(g() >> 11) * k53;
        const double r = std::sqrt(-2.0 * std::log(u1));
        re = r * std::cos(two_pi * u2);
        im = r * std::sin(two_pi * u2);
      }
      if (precision) {
        static_cast<double*>(host_out)[2 * i] = re;
        static_cast<double*>(host_out)[2 * i + 1] = im;
      } else {
        static_cast<float*>(host_out)[2 * i] = (float)re;
        static_cast<float*>(host_out)[2 * i + 1] = (float)im;
      }
    }
    return 0;
  });
}

int shafftAmdBatchPlan(shafft_amd_batch_plan** plan, int count, const int* ndims,
                       const int* dims, int precision) {
  return guarded([&] {
    if (!plan) fail(ERR_NULLPTR, "null plan pointer");
    *plan = nullptr;
    auto p = std::make_unique<shafft_amd_batch_plan>();
    p->be = std::make_unique<BatchPlan>(count, ndims, dims, precision);
    *plan = p.release();
    return 0;
  });
}

int shafftAmdBatchLayout(shafft_amd_batch_plan* plan, int64_t* total_elems, int64_t* offsets) {
  return guarded([&] {
    if (!plan || !total_elems) fail(ERR_NULLPTR, "null argument");
    *total_elems = plan->be->total_elems();
    if (offsets)
      for (int t = 0; t < plan->be->count(); ++t) offsets[t] = plan->be->offset(t);
    return 0;
  });
}

int shafftAmdBatchSetBuffers(shafft_amd_batch_plan* plan, void* in, void* out, void* work) {
  return guarded([&] {
    if (!plan) fail(ERR_NULLPTR, "null plan");
    plan->be->set_buffers(in, out, work);
    return 0;
  });
}

int shafftAmdBatchSetStream(shafft_amd_batch_plan* plan, void* stream) {
  return guarded([&] {
    if (!plan) fail(ERR_NULLPTR, "null plan");
    plan->be->set_stream(reinterpret_cast<hipStream_t>(stream));
    return 0;
  });
}

int shafftAmdBatchExecute(shafft_amd_batch_plan* plan, int direction, int use_graph) {
  return guarded([&] {
    if (!plan) fail(ERR_NULLPTR, "null plan");
    plan->be->execute(direction == 0 ? -1 : +1, use_graph != 0);
    return 0;
  });
}

int shafftAmdBatchSynchronize(shafft_amd_batch_plan* plan) {
  return guarded([&] {
    if (!plan) fail(ERR_NULLPTR, "null plan");
    plan->be->synchronize();
    return 0;
  });
}

int shafftAmdBatchSetPrecisionMode(shafft_amd_batch_plan* plan, int mode) {
  return guarded([&] {
    if (!plan) fail(ERR_NULLPTR, "null plan");
    plan->be->set_precision_mode(mode);
    return 0;
  });
}

int shafftAmdBatchAutotune(shafft_amd_batch_plan* plan, char* buf, int buflen) {
  return guarded([&] {
    if (!plan) fail(ERR_NULLPTR, "null plan");
    const std::string r = plan->be->autotune();
    if (buf && buflen > 0) copy_desc(r, buf, buflen);
    return 0;
  });
}

int shafftAmdBatchInfo(shafft_amd_batch_plan* plan, int* npasses, int64_t* nitems, int* grid,
                       int64_t* bytes) {
  return guarded([&] {
    if (!plan) fail(ERR_NULLPTR, "null plan");
    if (npasses) *npasses = plan->be->npasses();
    if (nitems) *nitems = plan->be->nitems();
    if (grid) *grid = plan->be->grid();
    if (bytes) *bytes = plan->be->total_bytes();
    return 0;
  });
}

int shafftAmdBatchDestroy(shafft_amd_batch_plan* plan) {
  return guarded([&] {
    delete plan;
    return 0;
  });
}

// Transform server -----------------------------------------------------------
int shafftAmdServerCreate(shafft_amd_server** srv, int slots, int64_t arena_bytes,
                          int planner_threads, int max_bypass) {
  return guarded([&] {
    if (!srv) fail(ERR_NULLPTR, "null server pointer");
    *srv = nullptr;
    Server::Config cfg;
    cfg.slots = slots;
    cfg.arena_bytes = arena_bytes;
    cfg.planners = planner_threads;
    cfg.max_bypass = max_bypass;
    auto p = std::make_unique<shafft_amd_server>();
    p->be = std::make_unique<Server>(cfg);
    *srv = p.release();
    return 0;
  });
}

int shafftAmdServerSubmitND(shafft_amd_server* srv, int ndim, const int* dims, int precision,
                            int direction, int normalize, const void* in, void* out,
                            int64_t arrival_ns, int64_t* job) {
  return guarded([&] {
    if (!srv || !dims || !job) fail(ERR_NULLPTR, "null argument");
    if (ndim < 1 || ndim > 8) fail(ERR_INVALID_DIM, "ndim must be in [1, 8]");
    ServeJob d;
    d.kind = 0;
    d.ndim = ndim;
    for (int i = 0; i < ndim; ++i) d.dims[i] = dims[i];
    d.prec = precision;
    d.dir = direction;
    d.normalize = normalize != 0;
    *job = srv->be->submit(d, in, out, arrival_ns);
    return 0;
  });
}

int shafftAmdServerSubmit1D(shafft_amd_server* srv, int64_t n, int precision, int direction,
                            int normalize, const void* in, void* out, int64_t arrival_ns,
                            int64_t* job) {
  return guarded([&] {
    if (!srv || !job) fail(ERR_NULLPTR, "null argument");
    ServeJob d;
    d.kind = 1;
    d.n = n;
    d.prec = precision;
    d.dir = direction;
    d.normalize = normalize != 0;
    *job = srv->be->submit(d, in, out, arrival_ns);
    return 0;
  });
}

int shafftAmdServerWait(shafft_amd_server* srv, int64_t job, int64_t timeout_ns, int* finished,
                        int* job_status, int64_t* times) {
  return guarded([&] {
    if (!srv || !finished) fail(ERR_NULLPTR, "null argument");
    int st = 0;
    ServeTimes t;
    std::string msg;
    *finished = srv->be->wait(job, timeout_ns, &st, &t, &msg) == 0;
    if (*finished) {
      if (job_status) *job_status = st;
      if (times) {
        times[0] = t.arrival;
        times[1] = t.admit;
        times[2] = t.done;
        times[3] = t.gpu_ns;
      }
      if (st) set_last_error(st, SRC_NONE, 0, msg.c_str());
    }
    return 0;
  });
}

int shafftAmdServerQuery(shafft_amd_server* srv, int64_t job, int* finished) {
  return guarded([&] {
    if (!srv || !finished) fail(ERR_NULLPTR, "null argument");
    *finished = srv->be->query(job) == 0;
    return 0;
  });
}

int shafftAmdServerDrain(shafft_amd_server* srv) {
  return guarded([&] {
    if (!srv) fail(ERR_NULLPTR, "null server");
    srv->be->drain();
    return 0;
  });
}

int shafftAmdServerStats(shafft_amd_server* srv, int64_t* counts, double* values) {
  return guarded([&] {
    if (!srv) fail(ERR_NULLPTR, "null server");
    const ServeStats s = srv->be->stats();
    if (counts) {
      const int64_t c[8] = {s.submitted, s.completed,   s.failed,      s.plans,
                            s.plan_builds, s.arena_bytes, s.arena_peak, s.table_bytes};
      for (int i = 0; i < 8; ++i) counts[i] = c[i];
    }
    if (values) {
      const double v[6] = {s.p50_ms, s.p99_ms, s.mean_ms, s.max_ms, s.span_s, s.gflop};
      for (int i = 0; i < 6; ++i) values[i] = v[i];
    }
    return 0;
  });
}

int shafftAmdServerSetBatching(shafft_amd_server* srv, int64_t max_elems, int max_jobs,
                               int max_inflight) {
  return guarded([&] {
    if (!srv) fail(ERR_NULLPTR, "null server");
    srv->be->set_batching(max_elems, max_jobs, max_inflight);
    return 0;
  });
}

int shafftAmdServerBatchStats(shafft_amd_server* srv, int64_t* batches, int64_t* batched_jobs) {
  return guarded([&] {
    if (!srv || !batches || !batched_jobs) fail(ERR_NULLPTR, "null argument");
    const ServeStats s = srv->be->stats();
    *batches = s.batches;
    *batched_jobs = s.batched_jobs;
    return 0;
  });
}

int shafftAmdServerResetStats(shafft_amd_server* srv) {
  return guarded([&] {
    if (!srv) fail(ERR_NULLPTR, "null server");
    srv->be->reset_stats();
    return 0;
  });
}

int shafftAmdServerNow(int64_t* ns) {
  return guarded([&] {
    if (!ns) fail(ERR_NULLPTR, "null argument");
    *ns = Server::now_ns();
    return 0;
  });
}

int shafftAmdServerDestroy(shafft_amd_server* srv) {
  return guarded([&] {
    delete srv;
    return 0;
  });
}

// =========================================================================
// 2. Reference C API at one rank
// =========================================================================

int shafftLastErrorStatus(void) { return last_error_status(); }
int shafftLastErrorSource(void) { return last_error_source(); }
int shafftLastErrorDomainCode(void) { return last_error_domain(); }
int shafftLastErrorMessage(char* buf, int buflen) {
  if (!buf || buflen <= 0) return (int)ERR_NULLPTR;
  const std::string& m = last_error_message();
  const size_t n = std::min<size_t>(m.size(), (size_t)buflen - 1);
  std::memcpy(buf, m.data(), n);
  buf[n] = 0;
  return (int)m.size();
}
void shafftClearLastError(void) { clear_last_error(); }
const char* shafftErrorSourceName(int source) {
  switch (source) {
    case SHAFFT_ERRSRC_NONE: return "NONE";
    case SHAFFT_ERRSRC_MPI: return "MPI";
    case SHAFFT_ERRSRC_HIP: return "HIP";
    case SHAFFT_ERRSRC_HIPFFT: return "hipFFT";
    case SHAFFT_ERRSRC_FFTW: return "FFTW";
    case SHAFFT_ERRSRC_SYSTEM: return "SYSTEM";
    default: return "UNKNOWN";
  }
}

int shafftNDCreate(void** outPlan) {
  return guarded([&] {
    if (!outPlan) fail(ERR_NULLPTR, "null outPlan");
    *outPlan = new Plan(Kind::ND);
    return 0;
  });
}

int shafft1DCreate(void** outPlan) {
  return guarded([&] {
    if (!outPlan) fail(ERR_NULLPTR, "null outPlan");
    *outPlan = new Plan(Kind::OneD);
    return 0;
  });
}

int shafftNDInit(void* planPtr, int ndim, int commDims[], int dimensions[], shafft_t precision,
                 shafft_comm_t cComm, shafft_transform_layout_t output_policy) {
  return guarded([&] {
    Plan* p = as_plan(planPtr);
    if (p->kind != Kind::ND) fail(ERR_INVALID_STATE, "not an N-D plan");
    if (p->state != State::UNINITIALIZED) fail(ERR_INVALID_STATE, "plan already initialised");
    const int prec = prec_of(precision);
    if (output_policy != SHAFFT_LAYOUT_REDISTRIBUTED && output_policy != SHAFFT_LAYOUT_INITIAL)
      fail(ERR_INVALID_LAYOUT, "invalid output policy");
    check_comm(cComm);
    if (!commDims || !dimensions) fail(ERR_NULLPTR, "null dims");
    if (ndim <= 0) fail(ERR_INVALID_DIM, "ndim must be positive");
    for (int i = 0; i < ndim; ++i)
      if (dimensions[i] <= 0) fail(ERR_INVALID_DIM, "dimension must be positive");
    if (commDims[ndim - 1] != 1) fail(ERR_INVALID_DECOMP, "last axis cannot be distributed");
    for (int i = 0; i < ndim; ++i) {
      if (commDims[i] < 1 || commDims[i] > dimensions[i])
        fail(ERR_INVALID_DECOMP, "invalid process grid");
      // One rank per plan: every grid extent must be 1.
      if (commDims[i] != 1) fail(ERR_INVALID_DECOMP, "process grid exceeds the communicator");
    }
    long double prod = 1.0L;
    size_t g = 1;
    for (int i = 0; i < ndim; ++i) {
      prod *= (long double)dimensions[i];
      g *= (size_t)dimensions[i];
    }
    p->prec = prec;
    p->dims.assign(dimensions, dimensions + ndim);
    p->global = g;
    p->norm_denominator = std::sqrt(prod);  // shafft_internal.cpp:329
    p->norm_exponent = 0;
    p->output_policy = output_policy;
    p->state = State::CONFIGURED;
    return 0;
  });
}

int shafft1DInit(void* planPtr, size_t N, size_t localN, size_t localStart, shafft_t precision,
                 shafft_comm_t cComm) {
  return guarded([&] {
    Plan* p = as_plan(planPtr);
    if (p->kind != Kind::OneD) fail(ERR_INVALID_STATE, "not a 1-D plan");
    if (p->state != State::UNINITIALIZED) fail(ERR_INVALID_STATE, "plan already initialised");
    const int prec = prec_of(precision);
    check_comm(cComm);
    if (N == 0) fail(ERR_INVALID_DIM, "N must be positive");
    if (localN != N || localStart != 0)
      fail(ERR_INVALID_LAYOUT, "layout does not match the one-rank distribution");
    p->prec = prec;
    p->dims = {(int)std::min<size_t>(N, INT_MAX)};
    p->global = N;
    p->norm_exponent = 0;
    p->layout_state = 0;
    p->state = State::CONFIGURED;
    return 0;
  });
}

int shafftConfiguration1D(size_t N, size_t* localN, size_t* localStart, size_t* localAllocSize,
                          shafft_t precision, shafft_comm_t cComm) {
  return guarded([&] {
    (void)prec_of(precision);
    check_comm(cComm);
    return shafftAmd1DQueryLayout(N, localN, localStart, localAllocSize, nullptr, nullptr);
  });
}

// One-rank outcome of the commDims -> nda -> strategy chain
// (src/cpp/configuration.cpp:530-600): a grid of ones (:478-498) or, for any
// other request, the automatic search, whose only candidate at one rank packs
// to d = 0 (:318-384) and so needs ndim >= 2.
int shafftConfigurationND(int ndim, int* size, shafft_t precision, int* commDims, int* nda,
                          size_t* subsize, size_t* offset, int* commSize,
                          shafft_decomposition_strategy_t strategy, size_t memLimit,
                          shafft_comm_t cComm) {
  (void)memLimit;
  return guarded([&] {
    (void)prec_of(precision);
    check_comm(cComm);
    if (!size || !commDims || !nda || !subsize || !offset || !commSize)
      fail(ERR_NULLPTR, "null argument");
    if (ndim <= 0) fail(ERR_INVALID_DIM, "ndim must be positive");
    if (strategy != SHAFFT_MAXIMIZE_NDA && strategy != SHAFFT_MINIMIZE_NDA)
      strategy = SHAFFT_MINIMIZE_NDA;  // shafftc.cpp:130-132
    for (int i = 0; i < ndim; ++i)
      if (size[i] <= 0) fail(ERR_INVALID_DIM, "dimension must be positive");
    bool ones = true;
    for (int i = 0; i < ndim; ++i) ones = ones && commDims[i] == 1;
    if (!ones && ndim < 2) fail(ERR_INVALID_DECOMP, "no decomposition of a 1-D tensor");
    for (int i = 0; i < ndim; ++i) {
      commDims[i] = 1;
      subsize[i] = (size_t)size[i];
      offset[i] = 0;
    }
    *nda = 0;
    *commSize = 1;
    return 0;
  });
}

// Fortran-facing companions (src/f03/shafftf03.cpp:11-60): the Fortran handle
// arrives by reference and is read as the one-rank handle value.
int shafftConfigurationNDf03(int ndim, int* size, shafft_t precision, int* commDims, int* nda,
                             size_t* subsize, size_t* offset, int* commSize,
                             shafft_decomposition_strategy_t strategy, size_t memLimit,
                             shafft_fint_t* fHandle) {
  return guarded([&]() -> int {
    if (!fHandle) fail(ERR_NULLPTR, "null communicator handle");
    return shafftConfigurationND(ndim, size, precision, commDims, nda, subsize, offset, commSize,
                                 strategy, memLimit, (shafft_comm_t)*fHandle);
  });
}

int shafftNDInitf03(void* planPtr, int ndim, int commDims[], int dimensions[],
                    shafft_t precision, shafft_fint_t* fHandle,
                    shafft_transform_layout_t* output_policy) {
  return guarded([&]() -> int {
    if (!fHandle || !output_policy) fail(ERR_NULLPTR, "null communicator handle or output policy");
    return shafftNDInit(planPtr, ndim, commDims, dimensions, precision, (shafft_comm_t)*fHandle,
                        *output_policy);
  });
}

int shafftConfiguration1Df03(size_t N, size_t* localN, size_t* localStart,
                             size_t* localAllocSize, shafft_t precision,
                             shafft_fint_t* fHandle) {
  return guarded([&]() -> int {
    if (!fHandle) fail(ERR_NULLPTR, "null communicator handle");
    return shafftConfiguration1D(N, localN, localStart, localAllocSize, precision,
                                 (shafft_comm_t)*fHandle);
  });
}

int shafft1DInitf03(void* planPtr, size_t N, size_t localN, size_t localStart,
                    shafft_t precision, shafft_fint_t* fHandle) {
  return guarded([&]() -> int {
    if (!fHandle) fail(ERR_NULLPTR, "null communicator handle");
    return shafft1DInit(planPtr, N, localN, localStart, precision, (shafft_comm_t)*fHandle);
  });
}

int shafftPlan(void* planPtr) {
  return guarded([&] {
    Plan* p = as_plan(planPtr);
    if (p->state == State::PLANNED) fail(ERR_INVALID_STATE, "plan() called twice");
    if (p->state != State::CONFIGURED) fail(ERR_PLAN_NOT_INIT, "plan not initialised");
    if (p->kind == Kind::ND) {
      const int ndim = (int)p->dims.size();
      std::vector<int> ta(ndim);
      for (int i = 0; i < ndim; ++i) ta[i] = i;
      p->nd = std::make_unique<NDBackend>(ndim, ta.data(), ndim, p->dims.data(), p->prec);
      p->nd->set_stream(p->stream);
    } else {
      p->od = std::make_unique<OneDBackend>((int64_t)p->global, p->prec);
      p->od->set_stream(p->stream);
    }
    p->state = State::PLANNED;
    return 0;
  });
}

int shafftDestroy(void** planPtr) {
  return guarded([&] {
    if (!planPtr || !*planPtr) fail(ERR_NULLPTR, "null plan");
    Plan* p = as_plan(*planPtr);
    p->magic = 0;
    delete p;
    *planPtr = nullptr;
    return 0;
  });
}

int shafftSetStream(void* planPtr, void* stream) {
  return guarded([&] {
    Plan* p = as_plan(planPtr);
    // N-D: the stream is applied to backend subplans, which exist only
    // after plan() (shafft_internal.cpp:198-210).
    if (p->kind == Kind::ND && p->state != State::PLANNED)
      fail(ERR_PLAN_NOT_INIT, "setStream before plan()");
    p->stream = reinterpret_cast<hipStream_t>(stream);
    if (p->nd) p->nd->set_stream(p->stream);
    if (p->od) p->od->set_stream(p->stream);
    return 0;
  });
}

int shafftGetLayout(void* planPtr, size_t* subsize, size_t* offset,
                    shafft_tensor_layout_t layout) {
  return guarded([&] {
    Plan* p = as_plan(planPtr);
    if (p->state == State::UNINITIALIZED) fail(ERR_PLAN_NOT_INIT, "plan not initialised");
    if (!subsize || !offset) fail(ERR_NULLPTR, "null output");
    if (layout < SHAFFT_TENSOR_LAYOUT_CURRENT || layout > SHAFFT_TENSOR_LAYOUT_REDISTRIBUTED)
      fail(ERR_INVALID_LAYOUT, "invalid layout");
    if (p->kind == Kind::ND) {
      for (size_t i = 0; i < p->dims.size(); ++i) {
        subsize[i] = (size_t)p->dims[i];
        offset[i] = 0;
      }
    } else {
      subsize[0] = p->global;
      offset[0] = 0;
    }
    return 0;
  });
}

int shafftGetAxes(void* planPtr, int* ca, int* da, shafft_tensor_layout_t layout) {
  return guarded([&] {
    Plan* p = as_plan(planPtr);
    if (p->state == State::UNINITIALIZED) fail(ERR_PLAN_NOT_INIT, "plan not initialised");
    if (layout < SHAFFT_TENSOR_LAYOUT_CURRENT || layout > SHAFFT_TENSOR_LAYOUT_REDISTRIBUTED)
      fail(ERR_INVALID_LAYOUT, "invalid layout");
    if (p->kind == Kind::ND) {
      if (!ca) fail(ERR_NULLPTR, "null ca");
      // One rank: every axis is contiguous, none distributed (slab_nd.cpp:522-526).
      for (size_t i = 0; i < p->dims.size(); ++i) ca[i] = (int)i;
    } else if (da) {
      // 1-D: no contiguous axis, the single axis is the distributed one
      // (src/cpp/shafft.cpp:776-786: nca 0, nda 1).
      da[0] = 0;
    }
    return 0;
  });
}

int shafftGetAllocSize(void* plan, size_t* localAllocSize) {
  return guarded([&] {
    Plan* p = as_plan(plan);
    if (!localAllocSize) fail(ERR_NULLPTR, "null output");
    if (p->state == State::UNINITIALIZED) fail(ERR_PLAN_NOT_INIT, "plan not initialised");
    *localAllocSize = p->global;
    return 0;
  });
}

int shafftGetGlobalSize(void* plan, size_t* globalSize) {
  return guarded([&] {
    Plan* p = as_plan(plan);
    if (!globalSize) fail(ERR_NULLPTR, "null output");
    if (p->state == State::UNINITIALIZED) fail(ERR_PLAN_NOT_INIT, "plan not initialised");
    *globalSize = p->global;
    return 0;
  });
}

int shafftIsConfigured(void* plan, int* configured) {
  return guarded([&] {
    if (!configured) fail(ERR_NULLPTR, "null output");
    *configured = 0;
    if (plan) *configured = as_plan(plan)->state != State::UNINITIALIZED ? 1 : 0;
    return 0;
  });
}

int shafftIsActive(void* plan, int* active) {
  return guarded([&] {
    if (!active) fail(ERR_NULLPTR, "null output");
    *active = 0;
    if (plan) *active = as_plan(plan)->state != State::UNINITIALIZED ? 1 : 0;
    return 0;
  });
}

int shafftSetBuffers(void* planPtr, void* data, void* work) {
  return guarded([&] {
    Plan* p = as_plan(planPtr);
    if (p->state == State::UNINITIALIZED) fail(ERR_PLAN_NOT_INIT, "plan not initialised");
    if (!data || !work) fail(ERR_NULLPTR, "null buffer");
    p->data = data;
    p->work = work;
    return 0;
  });
}

int shafftGetBuffers(void* planPtr, void** data, void** work) {
  return guarded([&] {
    Plan* p = as_plan(planPtr);
    if (!data || !work) fail(ERR_NULLPTR, "null output");
    if (p->state == State::UNINITIALIZED) fail(ERR_PLAN_NOT_INIT, "plan not initialised");
    *data = p->data;
    *work = p->work;
    return 0;
  });
}

int shafftExecute(void* planPtr, shafft_direction_t direction) {
  return guarded([&] {
    Plan* p = as_plan(planPtr);
    if (p->state != State::PLANNED) fail(ERR_PLAN_NOT_INIT, "plan() not called");
    if (direction != SHAFFT_FORWARD && direction != SHAFFT_BACKWARD)
      fail(ERR_INVALID_STATE, "invalid direction");
    if (!p->data || !p->work) fail(ERR_NO_BUFFER, "buffers not set");
    const int dir = direction == SHAFFT_FORWARD ? -1 : +1;
    if (p->kind == Kind::ND) {
      p->nd->execute(p->data, p->work, dir);
    } else {
      // data -> work, then swap so data holds the result
      // (shafft_internal.cpp:596-611); host-synchronous like the N-D path.
      p->od->execute(p->data, p->work, dir);
      hip_check(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
      if (p->work != p->data) std::swap(p->data, p->work);
      p->layout_state = direction == SHAFFT_FORWARD ? 1 : 0;
    }
    p->norm_exponent += 1;
    return 0;
  });
}

int shafftNormalize(void* planPtr) {
  return guarded([&] {
    Plan* p = as_plan(planPtr);
    if (p->state == State::UNINITIALIZED) fail(ERR_PLAN_NOT_INIT, "plan not initialised");
    if (!p->data || !p->work) fail(ERR_NO_BUFFER, "buffers not set");
    if (p->kind == Kind::ND) {
      // factor = 1 / denominator^e in long double, rounded once to the
      // working precision (shafft_internal.cpp:136,150-156).
      long double f = 1.0 / std::pow(p->norm_denominator, p->norm_exponent);
      const double fd = p->prec ? (double)f : (double)(float)f;
      scale_async(p->data, fd, p->global, p->prec, p->stream);
      hip_check(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
    } else {
      if (!p->od) fail(ERR_PLAN_NOT_INIT, "plan() not called");
      const size_t n = p->global;
      double f;
      if (p->prec == 0) {
        float norm = 1.0f / std::pow(std::sqrt(static_cast<float>(n)), p->norm_exponent);
        f = norm;
      } else {
        f = 1.0 / std::pow(std::sqrt(static_cast<double>(n)), p->norm_exponent);
      }
      scale_async(p->data, f, n, p->prec, p->stream);
      hip_check(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
    }
    p->norm_exponent = 0;
    return 0;
  });
}

// Device allocation of `count` elements of `esz` bytes, as
// src/cpp/detail/buffer_utils.hpp:37-56: count 0 gives a null buffer and
// success, a failed (or unrepresentable) allocation ERR_ALLOC.
static int alloc_elems(size_t count, size_t esz, void** buf) {
  return guarded([&] {
    if (!buf) fail(ERR_NULLPTR, "null output");
    *buf = nullptr;
    if (count == 0) return 0;
    if (count > SIZE_MAX / esz) fail(ERR_ALLOC, "allocation size overflows size_t");
    const hipError_t e = hipMalloc(buf, count * esz);
    if (e != hipSuccess) {
      *buf = nullptr;
      throw Error(ERR_ALLOC, SRC_HIP, static_cast<int>(e),
                  std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    return 0;
  });
}

int shafftAllocBufferF(size_t count, void** buf) { return alloc_elems(count, 8, buf); }

int shafftAllocBufferD(size_t count, void** buf) { return alloc_elems(count, 16, buf); }

int shafftFreeBufferF(void* buf) {
  return guarded([&] {
    if (buf) hip_check(hipFree(buf), "hipFree");
    return 0;
  });
}

int shafftFreeBufferD(void* buf) { return shafftFreeBufferF(buf); }

static int copy_bytes(void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  return guarded([&] {
    if (!dst || !src) fail(ERR_NULLPTR, "null pointer");
    if (bytes == 0) return 0;
    hip_check(hipMemcpy(dst, src, bytes, kind), "hipMemcpy");
    return 0;
  });
}

int shafftCopyToBufferF(void* dst, const void* src, size_t count) {
  return copy_bytes(dst, src, count * 8, hipMemcpyHostToDevice);
}
int shafftCopyToBufferD(void* dst, const void* src, size_t count) {
  return copy_bytes(dst, src, count * 16, hipMemcpyHostToDevice);
}
int shafftCopyFromBufferF(void* dst, const void* src, size_t count) {
  return copy_bytes(dst, src, count * 8, hipMemcpyDeviceToHost);
}
int shafftCopyFromBufferD(void* dst, const void* src, size_t count) {
  return copy_bytes(dst, src, count * 16, hipMemcpyDeviceToHost);
}

const char* shafftGetBackendName(void) { return "gfx950"; }

void shafftGetVersion(int* major, int* minor, int* patch) {
  if (major) *major = SHAFFT_AMD_VERSION_MAJOR;
  if (minor) *minor = SHAFFT_AMD_VERSION_MINOR;
  if (patch) *patch = SHAFFT_AMD_VERSION_PATCH;
}

const char* shafftGetVersionString(void) { return "1.1.0"; }

int shafftFinalize(void) { return 0; }

// ---------------------------------------------------------------------------
// Configuration objects at one rank (shafft.h:203-475; resolution rules of
// src/cpp/config_info.cpp:200-510 with a communicator of size 1).
// ---------------------------------------------------------------------------

static const int kConfigApiVersion = 1;

static char* dup_cstr(const char* src, size_t& len) {
  len = src ? std::strlen(src) : 0;
  if (!src) return nullptr;
  char* d = static_cast<char*>(std::malloc(len + 1));
  if (!d) fail(ERR_ALLOC, "config string");
  std::memcpy(d, src, len + 1);
  return d;
}

extern "C++" {
// Host name and current HIP device name (none: NULL), config-owned.
template <typename C>
static void fill_topology(C* cfg) {
  char host[256] = {0};
  if (gethostname(host, sizeof host - 1) != 0) std::strcpy(host, "localhost");
  std::free(cfg->hostname);
  cfg->hostname = dup_cstr(host, cfg->hostnameLen);
  std::free(cfg->deviceName);
  cfg->deviceName = nullptr;
  cfg->deviceNameLen = 0;
  int n = 0, dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDeviceCount(&n) == hipSuccess && n > 0 && hipGetDevice(&dev) == hipSuccess &&
      hipGetDeviceProperties(&prop, dev) == hipSuccess)
    cfg->deviceName = dup_cstr(prop.gcnArchName[0] ? prop.gcnArchName : prop.name, cfg->deviceNameLen);
  (void)hipGetLastError();
  cfg->nodeId = 0;
  cfg->nodeCount = 1;
}

template <typename T>
static T* alloc_array(int n) {
  T* a = static_cast<T*>(std::calloc((size_t)n, sizeof(T)));
  if (!a) fail(ERR_ALLOC, "config array");
  return a;
}
}  // extern "C++"

static void nd_config_free(shafft_nd_config_t* c) {
  void* owned[] = {c->globalShape,          c->hintCommDims,           c->commDims,
                   c->initial.subsize,      c->initial.offset,         c->initial.contiguousAxes,
                   c->initial.distributedAxes, c->output.subsize,      c->output.offset,
                   c->output.contiguousAxes, c->output.distributedAxes, c->hostname,
                   c->deviceName};
  for (void* q : owned) std::free(q);
  std::memset(c, 0, sizeof *c);
}

void shafftConfigNDRelease(shafft_nd_config_t* cfg) {
  if (!cfg || cfg->structSize == 0) return;  // idempotent; safe after a failed Init
  nd_config_free(cfg);
}

int shafftConfigNDResolve(shafft_nd_config_t* cfg) {
  return guarded([&] {
    if (!cfg) fail(ERR_NULLPTR, "null config");
    if (cfg->structSize != sizeof(shafft_nd_config_t)) fail(ERR_INVALID_STATE, "config not initialised");
    if (cfg->ndim < 1) fail(ERR_INVALID_DIM, "ndim must be positive");
    if (cfg->worldComm != 0) fail(ERR_INVALID_COMM, "communicator must be the one-rank handle 0");
    const int nd = cfg->ndim;
    cfg->flags &= ~SHAFFT_CONFIG_RESOLVED;
    try {
      for (int i = 0; i < nd; ++i) {
        if (cfg->globalShape[i] == 0) fail(ERR_INVALID_DIM, "extent must be positive");
        if (cfg->globalShape[i] > (size_t)INT_MAX) fail(ERR_SIZE_OVERFLOW, "extent exceeds INT_MAX");
      }
      if (cfg->precision != SHAFFT_C2C && cfg->precision != SHAFFT_Z2Z)
        fail(ERR_INVALID_FFTTYPE, "invalid precision");
    } catch (const Error& e) {
      cfg->status = e.status;
      throw;
    }
    // One rank: every hint resolves to the grid of ones and no distributed axis.
    bool grid_changed = false;
    size_t total = 1;
    for (int i = 0; i < nd; ++i) {
      cfg->commDims[i] = 1;
      grid_changed = grid_changed || cfg->hintCommDims[i] != 1;
      total *= cfg->globalShape[i];
    }
    cfg->nda = 0;
    cfg->commSize = 1;
    cfg->allocElements = total;
    for (shafft_nd_layout_t* l : {&cfg->initial, &cfg->output}) {
      for (int i = 0; i < nd; ++i) {
        l->subsize[i] = cfg->globalShape[i];
        l->offset[i] = 0;
        l->contiguousAxes[i] = i;
        l->distributedAxes[i] = 0;  // length nda = 0: unused
      }
      l->localElements = total;
      l->hasLocalElements = total > 0;
    }
    cfg->isActive = 1;
    cfg->activeRank = 0;
    cfg->activeSize = 1;
    cfg->activeComm = 0;
    fill_topology(cfg);
    cfg->flags = SHAFFT_CONFIG_RESOLVED;
    if (grid_changed) cfg->flags |= SHAFFT_CONFIG_CHANGED_COMM_DIMS;
    if (cfg->hintNda != cfg->nda) cfg->flags |= SHAFFT_CONFIG_CHANGED_NDA;
    cfg->status = 0;
    return 0;
  });
}

int shafftConfigNDInit(shafft_nd_config_t* cfg, int ndim, const size_t* globalShape,
                       shafft_t precision, const int* commDims, int hintNda,
                       shafft_decomposition_strategy_t strategy,
                       shafft_transform_layout_t outputPolicy, size_t memLimit,
                       shafft_comm_t cComm) {
  const int rc = guarded([&] {
    if (!cfg) fail(ERR_NULLPTR, "null config");
    if (ndim < 1) fail(ERR_INVALID_DIM, "ndim must be positive");
    if (!globalShape) fail(ERR_NULLPTR, "null globalShape");
    if (cfg->structSize == sizeof(shafft_nd_config_t)) fail(ERR_INVALID_STATE, "config already initialised");
    if (cComm != 0) fail(ERR_INVALID_COMM, "communicator must be the one-rank handle 0");
    if (outputPolicy != SHAFFT_LAYOUT_REDISTRIBUTED && outputPolicy != SHAFFT_LAYOUT_INITIAL)
      fail(ERR_INVALID_LAYOUT, "invalid output policy");
    std::memset(cfg, 0, sizeof *cfg);
    cfg->structSize = sizeof(shafft_nd_config_t);
    cfg->apiVersion = kConfigApiVersion;
    cfg->ndim = ndim;
    cfg->activeRank = -1;
    try {
      cfg->globalShape = alloc_array<size_t>(ndim);
      cfg->hintCommDims = alloc_array<int>(ndim);
      cfg->commDims = alloc_array<int>(ndim);
      for (shafft_nd_layout_t* l : {&cfg->initial, &cfg->output}) {
        l->subsize = alloc_array<size_t>(ndim);
        l->offset = alloc_array<size_t>(ndim);
        l->contiguousAxes = alloc_array<int>(ndim);
        l->distributedAxes = alloc_array<int>(ndim);
      }
    } catch (...) {
      nd_config_free(cfg);
      throw;
    }
    for (int i = 0; i < ndim; ++i) {
      cfg->globalShape[i] = globalShape[i];
      cfg->hintCommDims[i] = commDims ? commDims[i] : 0;  // 0: no preference
    }
    cfg->precision = precision;
    cfg->hintNda = hintNda;
    cfg->strategy = strategy;
    cfg->outputPolicy = outputPolicy;
    cfg->memLimit = memLimit;
    cfg->worldComm = 0;
    cfg->activeComm = 0;
    return 0;
  });
  if (rc != 0) return rc;
  const int r2 = shafftConfigNDResolve(cfg);
  if (r2 != 0) nd_config_free(cfg);  // as a failed Init: nothing left to release
  return r2;
}

int shafftNDInitFromConfig(void* planPtr, shafft_nd_config_t* cfg) {
  if (!cfg) return guarded([&]() -> int { fail(ERR_NULLPTR, "null config"); });
  if (!(cfg->flags & SHAFFT_CONFIG_RESOLVED)) {
    const int rc = shafftConfigNDResolve(cfg);
    if (rc != 0) return rc;
  }
  std::vector<int> dims(cfg->ndim), grid(cfg->ndim);
  for (int i = 0; i < cfg->ndim; ++i) {
    dims[i] = (int)cfg->globalShape[i];
    grid[i] = cfg->commDims[i];
  }
  return shafftNDInit(planPtr, cfg->ndim, grid.data(), dims.data(), cfg->precision,
                      cfg->worldComm, cfg->outputPolicy);
}

void shafftConfig1DRelease(shafft_1d_config_t* cfg) {
  if (!cfg || cfg->structSize == 0) return;
  std::free(cfg->hostname);
  std::free(cfg->deviceName);
  std::memset(cfg, 0, sizeof *cfg);
}

int shafftConfig1DResolve(shafft_1d_config_t* cfg) {
  return guarded([&] {
    if (!cfg) fail(ERR_NULLPTR, "null config");
    if (cfg->structSize != sizeof(shafft_1d_config_t)) fail(ERR_INVALID_STATE, "config not initialised");
    if (cfg->globalSize == 0) fail(ERR_INVALID_DIM, "globalSize must be positive");
    if (cfg->worldComm != 0) fail(ERR_INVALID_COMM, "communicator must be the one-rank handle 0");
    cfg->flags &= ~SHAFFT_CONFIG_RESOLVED;
    for (shafft_1d_layout_t* l : {&cfg->initial, &cfg->output}) {
      l->localSize = cfg->globalSize;
      l->localStart = 0;
      l->hasLocalElements = 1;
    }
    cfg->allocElements = cfg->globalSize;
    cfg->isActive = 1;
    cfg->activeRank = 0;
    cfg->activeSize = 1;
    cfg->activeComm = 0;
    fill_topology(cfg);
    cfg->flags = SHAFFT_CONFIG_RESOLVED;
    cfg->status = 0;
    return 0;
  });
}

int shafftConfig1DInit(shafft_1d_config_t* cfg, size_t globalSize, shafft_t precision,
                       shafft_comm_t cComm) {
  const int rc = guarded([&] {
    if (!cfg) fail(ERR_NULLPTR, "null config");
    if (globalSize == 0) fail(ERR_INVALID_DIM, "globalSize must be positive");
    if (precision != SHAFFT_C2C && precision != SHAFFT_Z2Z) fail(ERR_INVALID_FFTTYPE, "invalid precision");
    if (cfg->structSize == sizeof(shafft_1d_config_t)) fail(ERR_INVALID_STATE, "config already initialised");
    if (cComm != 0) fail(ERR_INVALID_COMM, "communicator must be the one-rank handle 0");
    std::memset(cfg, 0, sizeof *cfg);
    cfg->structSize = sizeof(shafft_1d_config_t);
    cfg->apiVersion = kConfigApiVersion;
    cfg->globalSize = globalSize;
    cfg->precision = precision;
    cfg->activeRank = -1;
    return 0;
  });
  if (rc != 0) return rc;
  const int r2 = shafftConfig1DResolve(cfg);
  if (r2 != 0) shafftConfig1DRelease(cfg);
  return r2;
}

int shafft1DInitFromConfig(void* planPtr, shafft_1d_config_t* cfg) {
  if (!cfg) return guarded([&]() -> int { fail(ERR_NULLPTR, "null config"); });
  if (!(cfg->flags & SHAFFT_CONFIG_RESOLVED)) {
    const int rc = shafftConfig1DResolve(cfg);
    if (rc != 0) return rc;
  }
  return shafft1DInit(planPtr, cfg->globalSize, cfg->initial.localSize, cfg->initial.localStart,
                      cfg->precision, cfg->worldComm);
}

int shafftGetCommunicator(void* planPtr, shafft_comm_t* outComm) {
  return guarded([&] {
    if (!outComm) fail(ERR_NULLPTR, "null output");
    Plan* p = as_plan(planPtr);
    if (p->state == State::UNINITIALIZED) fail(ERR_PLAN_NOT_INIT, "plan not initialised");
    *outComm = 0;
    return 0;
  });
}

}  // extern "C"
