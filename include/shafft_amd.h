/* shafft_amd.h - C ABI of the gfx950 (MI355X) slab C2C FFT backend.
 *
 * Two groups of entry points, all extern "C", plain pointers and sizes:
 *
 *  1. The backend seam.  SHAFFT's core calls its FFT backend through a small
 *     link-time symbol set (fftndPlan/fftndExecute/..., fft1d*, normalize*).
 *     Each shafftAmd* function below replaces one of them; the C++ adapter
 *     that binds them to the reference's FFTNDHandle / FFT1DHandle types
 *     forwards each seam call here.  Plans are opaque handles owned by the
 *     library; data/work buffers are owned by the caller.
 *
 *  2. The reference's public C API (include/shafft/shafft.h) at one rank per
 *     GPU: the same function names, argument order, enum values and status
 *     codes.  The MPI communicator argument is kept in its position as an
 *     opaque handle that must describe a one-rank communicator (pass 0); the
 *     process grid is fixed to {1,...,1}.
 *
 * Directions: 0 = forward (exp(-2 pi i k.x / N)), 1 = backward, both
 * unnormalised.  Precision: 0 = complex float (C2C), 1 = complex double (Z2Z).
 * Tensors are row-major, last axis fastest.  Every status is 0 on success or
 * a shafft_status_t value; the thread-local last-error record gives details.
 */
#ifndef SHAFFT_AMD_H
#define SHAFFT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ======================================================================== */
/* 1. Backend seam                                                          */
/* ======================================================================== */

typedef struct shafft_amd_nd_plan shafft_amd_nd_plan;
typedef struct shafft_amd_1d_plan shafft_amd_1d_plan;

/* Replaces fftndPlan (src/cpp/fft/hipfft_method/fftnd_method.hpp:14-21).
 * Local N-D transform over the consecutive axes ta[0..nta-1] of the
 * row-major tensor dimensions[0..ndim-1].  in/out may be NULL (as in the
 * reference: tests/unit/test_fft_partial.cpp:151-158). */
int shafftAmdNDPlan(shafft_amd_nd_plan** plan, int nta, const int* ta, int ndim,
                    const int* dimensions, int precision, void* in, void* out);

/* Replaces fftndSetStream (fftnd_method.hpp:23).  stream is a hipStream_t. */
int shafftAmdNDSetStream(shafft_amd_nd_plan* plan, void* stream);

/* Replaces fftndExecute (fftnd_method.hpp:25): reads *data, may overwrite
 * *work and *data, and on return *data holds the result (the two pointers
 * are swapped when needed, as in fftnd_method.cpp:514-526).  Returns after
 * the stream has drained (fftnd_method.cpp:520). */
int shafftAmdNDExecute(shafft_amd_nd_plan* plan, void** data, void** work, int direction);

/* Replaces fftndDestroy (fftnd_method.hpp:27). */
int shafftAmdNDDestroy(shafft_amd_nd_plan* plan);

/* Replaces fftndFinalize (fftnd_method.hpp:29-31). */
int shafftAmdNDFinalize(void);

/* Replace normalizeComplexFloat / normalizeComplexDouble
 * (src/cpp/fft/hipfft_method/common/normalize.hpp): x[i] *= factor over
 * tensor_size complex elements, asynchronous on stream. */
int shafftAmdNormalizeComplexFloat(float factor, size_t tensor_size, void* data, void* stream);
int shafftAmdNormalizeComplexDouble(double factor, size_t tensor_size, void* data, void* stream);

/* Replaces shafft::fft1dQueryLayout (fft1d_method.hpp:34-41) for one rank. */
int shafftAmd1DQueryLayout(size_t globalN, size_t* localN, size_t* localStart,
                           size_t* localAllocSize, size_t* localNTrans,
                           size_t* localStartTrans);

/* Replaces shafft::fft1dPlan (fft1d_method.hpp:16-22) for one rank. */
int shafftAmd1DPlan(shafft_amd_1d_plan** plan, size_t globalN, int precision, void* in,
                    void* out);

/* Replaces shafft::fft1dExecute (fft1d_method.hpp:24): in -> out, input
 * preserved unless in == out.  Asynchronous on the plan's stream. */
int shafftAmd1DExecute(shafft_amd_1d_plan* plan, void* in, void* out, int direction);

/* Allocation-free variant of shafftAmd1DExecute, for callers that cannot
 * allocate device memory while they run (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  No reference
 * counterpart.  Same transform, arguments and kernels as shafftAmd1DExecute,
 * but the plan's intermediates (the Bluestein convolution, the in-place and
 * input-preserving four-step temporaries) go to one library-owned device
 * array of 64 MiB per GPU (held until the process ends) instead of plan
 * scratch, and the call allocates nothing.  A plan needing more fails with
 * SHAFFT_ERR_ALLOC (e.g. 1,000,003 points fit in both precisions; 2^27
 * points do not).  Every call on a GPU shares that array: calls that
 * overlap on one GPU race.  Asynchronous on the plan's stream. */
int shafftAmd1DExecuteNoAlloc(shafft_amd_1d_plan* plan, const void* in, void* out,
                              int direction);

/* Replace shafft::fft1dSetStream / fft1dSynchronize / fft1dDestroy. */
int shafftAmd1DSetStream(shafft_amd_1d_plan* plan, void* stream);
int shafftAmd1DSynchronize(shafft_amd_1d_plan* plan);
int shafftAmd1DDestroy(shafft_amd_1d_plan* plan);

/* Replaces shafft::fft1dNormalize (fft1d_method.hpp:32): factor
 * 1/sqrt(N)^normExponent in the working precision, then synchronise. */
int shafftAmd1DNormalize(shafft_amd_1d_plan* plan, void* data, int normExponent,
                         size_t localCount);

/* Introspection used by the benchmark and profiles (no reference
 * counterpart): number of LDS passes, algorithmic HBM bytes of pass i, the
 * element counts pass i reads and writes (a Bluestein chain's intermediate
 * passes span the padded length), and a single-pass launch in -> out
 * (asynchronous on the plan's stream) that refuses buffers smaller than
 * those extents. */
int shafftAmdNDNumPasses(shafft_amd_nd_plan* plan, int* npasses);
int shafftAmdNDPassBytes(shafft_amd_nd_plan* plan, int pass, int64_t* bytes);
int shafftAmdNDPassExtents(shafft_amd_nd_plan* plan, int pass, int64_t* in_elems,
                           int64_t* out_elems);
int shafftAmdNDExecutePass(shafft_amd_nd_plan* plan, int pass, void* in, int64_t in_elems,
                           void* out, int64_t out_elems, int direction);
int shafftAmd1DNumPasses(shafft_amd_1d_plan* plan, int* npasses);
int shafftAmd1DPassBytes(shafft_amd_1d_plan* plan, int pass, int64_t* bytes);
int shafftAmd1DPassExtents(shafft_amd_1d_plan* plan, int pass, int64_t* in_elems,
                           int64_t* out_elems);
int shafftAmd1DExecutePass(shafft_amd_1d_plan* plan, int pass, void* in, int64_t in_elems,
                           void* out, int64_t out_elems, int direction);

/* Human-readable pass list of a plan into buf (NUL-terminated, truncated). */
int shafftAmdNDDescribe(shafft_amd_nd_plan* plan, char* buf, int buflen);
int shafftAmd1DDescribe(shafft_amd_1d_plan* plan, char* buf, int buflen);

/* Synthetic input of the benchmark workloads into host memory: `count`
 * interleaved complex values (precision 0 float, 1 double) from
 * std::mt19937_64(seed); distribution 0 = re, im uniform in [-1, 1) as
 * 2 ((x >> 11) 2^-53) - 1, 1 = Box-Muller Gaussian per element.  No reference
 * counterpart (the reference's tests draw their inputs from std::mt19937_64,
 * tests/integration/cpp/test_correctness.cpp:382). */
int shafftAmdSyntheticInput(void* host_out, int64_t count, uint64_t seed, int distribution,
                            int precision);

/* Batched uneven transforms (no reference counterpart: the reference runs
 * one plan per shape).  `count` independent transforms, transform t having
 * ndims[t] axes taken in order from dims (row-major, last axis fastest),
 * packed back to back (offsets from shafftAmdBatchLayout) in one buffer of
 * one precision.  Every axis must fit one LDS pass (length <= 8192 complex
 * float / 4096 complex double, or a Bluestein length whose padded
 * convolution does).  All passes of all transforms run as one persistent
 * launch; execute replays a captured HIP graph per direction. */
typedef struct shafft_amd_batch_plan shafft_amd_batch_plan;
int shafftAmdBatchPlan(shafft_amd_batch_plan** plan, int count, const int* ndims,
                       const int* dims, int precision);
int shafftAmdBatchLayout(shafft_amd_batch_plan* plan, int64_t* total_elems, int64_t* offsets);
int shafftAmdBatchSetBuffers(shafft_amd_batch_plan* plan, void* in, void* out, void* work);
int shafftAmdBatchSetStream(shafft_amd_batch_plan* plan, void* stream);
/* Asynchronous on the plan's stream; use_graph = 0 launches directly. */
int shafftAmdBatchExecute(shafft_amd_batch_plan* plan, int direction, int use_graph);
int shafftAmdBatchSynchronize(shafft_amd_batch_plan* plan);
/* Radix-8/16 stages of a complex-float batch: 0 FP32 (default), 1 bfloat16
 * hi/lo split precision on the matrix cores (FP32 accumulation), 2 float16
 * hi/lo split with per-butterfly power-of-two scaling.  Other radices stay
 * FP32.  No reference counterpart (the reference computes in FP32/FP64). */
int shafftAmdBatchSetPrecisionMode(shafft_amd_batch_plan* plan, int mode);
/* Times the candidate tile schedules on the attached buffers (overwrites
 * out and work), keeps the fastest and writes a report into buf. */
int shafftAmdBatchAutotune(shafft_amd_batch_plan* plan, char* buf, int buflen);
int shafftAmdBatchInfo(shafft_amd_batch_plan* plan, int* npasses, int64_t* nitems, int* grid,
                       int64_t* bytes);
int shafftAmdBatchDestroy(shafft_amd_batch_plan* plan);

/* Transform server (no reference counterpart: the reference runs one
 * synchronous transform per call, shafft_internal.cpp:62-127).  A request
 * loop that admits FFTND (all axes, row-major dims) and FFT1D jobs of mixed
 * shape and precision onto `slots` device streams as slots and scratch free
 * up.  Scratch comes from one device arena of `arena_bytes` allocated at
 * creation; a job waits until its scratch footprint fits (jobs whose
 * footprint exceeds the arena fail with SHAFFT_ERR_ALLOC), and a later job
 * that fits may pass an earlier waiting one at most `max_bypass` times.
 * Plans are built on `planner_threads` host threads and stay resident.
 * Jobs read `in` (left unchanged) and write `out`: caller-owned device
 * buffers of the job's element count, distinct, valid until the job ends.
 * direction: 0 forward, 1 backward; normalize != 0 scales the output by
 * 1/sqrt(elements).  arrival_ns < 0 stamps the submit time; otherwise it is
 * a shafftAmdServerNow() time from which latency is counted.  One slot per
 * hardware queue the HIP runtime gives the process (GPU_MAX_HW_QUEUES, 4 by
 * default) serves best: streams beyond that share queues and serialise. */
typedef struct shafft_amd_server shafft_amd_server;
int shafftAmdServerCreate(shafft_amd_server** srv, int slots, int64_t arena_bytes,
                          int planner_threads, int max_bypass);
int shafftAmdServerSubmitND(shafft_amd_server* srv, int ndim, const int* dims, int precision,
                            int direction, int normalize, const void* in, void* out,
                            int64_t arrival_ns, int64_t* job);
int shafftAmdServerSubmit1D(shafft_amd_server* srv, int64_t n, int precision, int direction,
                            int normalize, const void* in, void* out, int64_t arrival_ns,
                            int64_t* job);
/* Waits up to timeout_ns (< 0: no limit).  *finished = 1 when the job has
 * ended: its status goes to *job_status (0 = success; otherwise the job's
 * message is the calling thread's last error), times[4] = arrival, admit,
 * completion (steady-clock ns) and device ns, and the id is released. */
int shafftAmdServerWait(shafft_amd_server* srv, int64_t job, int64_t timeout_ns, int* finished,
                        int* job_status, int64_t* times);
int shafftAmdServerQuery(shafft_amd_server* srv, int64_t job, int* finished);
int shafftAmdServerDrain(shafft_amd_server* srv);
/* counts[8] = submitted, completed, failed, plans, plan builds, arena bytes,
 * arena peak bytes, resident table bytes; values[6] = p50, p99, mean, max
 * latency (ms), span from first arrival to last completion (s), GFLOP
 * completed (5 N log2 N). */
int shafftAmdServerStats(shafft_amd_server* srv, int64_t* counts, double* values);
int shafftAmdServerResetStats(shafft_amd_server* srv);
/* Micro-batching of small jobs (default: up to 2^20 elements, 64 jobs a
 * group, 2 groups in flight per precision): jobs of at most max_elems
 * elements (0: off) that are admitted in one scheduling round are grouped
 * per precision, up to max_jobs a group, and each group runs as one
 * persistent batch-kernel launch per pass level on one slot; while
 * max_inflight groups of a precision run, its small jobs wait and gather
 * into the next group.  A grouped job's result is the same alone and in any
 * mix and equals the plain call's within the precision's tolerance (jobs
 * with normalize != 0 and plans with the fixed-length Bluestein kernels
 * always run their own launches).  Applies to jobs admitted after the call. */
int shafftAmdServerSetBatching(shafft_amd_server* srv, int64_t max_elems, int max_jobs,
                               int max_inflight);
/* Micro-batches formed and the jobs they held since creation or the last
 * shafftAmdServerResetStats. */
int shafftAmdServerBatchStats(shafft_amd_server* srv, int64_t* batches, int64_t* batched_jobs);
int shafftAmdServerNow(int64_t* ns);
int shafftAmdServerDestroy(shafft_amd_server* srv); /* drains first */

/* ======================================================================== */
/* 2. Reference C API (include/shafft/shafft.h) at one rank per GPU          */
/* ======================================================================== */

typedef enum { SHAFFT_C2C, SHAFFT_Z2Z } shafft_t;
typedef enum { SHAFFT_FORWARD, SHAFFT_BACKWARD } shafft_direction_t;
typedef enum {
  SHAFFT_TENSOR_LAYOUT_CURRENT,
  SHAFFT_TENSOR_LAYOUT_INITIAL,
  SHAFFT_TENSOR_LAYOUT_REDISTRIBUTED
} shafft_tensor_layout_t;
typedef enum { SHAFFT_LAYOUT_REDISTRIBUTED = 0, SHAFFT_LAYOUT_INITIAL = 1 } shafft_transform_layout_t;
typedef enum {
  SHAFFT_SUCCESS = 0,
  SHAFFT_ERR_NULLPTR = 1,
  SHAFFT_ERR_INVALID_COMM = 2,
  SHAFFT_ERR_NO_BUFFER = 3,
  SHAFFT_ERR_PLAN_NOT_INIT = 4,
  SHAFFT_ERR_INVALID_DIM = 5,
  SHAFFT_ERR_DIM_MISMATCH = 6,
  SHAFFT_ERR_INVALID_DECOMP = 7,
  SHAFFT_ERR_INVALID_FFTTYPE = 8,
  SHAFFT_ERR_ALLOC = 9,
  SHAFFT_ERR_BACKEND = 10,
  SHAFFT_ERR_MPI = 11,
  SHAFFT_ERR_INVALID_LAYOUT = 12,
  SHAFFT_ERR_SIZE_OVERFLOW = 13,
  SHAFFT_ERR_NOT_IMPL = 14,
  SHAFFT_ERR_INVALID_STATE = 15,
  SHAFFT_ERR_INTERNAL = 16
} shafft_status_t;
typedef enum {
  SHAFFT_ERRSRC_NONE = 0,
  SHAFFT_ERRSRC_MPI = 1,
  SHAFFT_ERRSRC_HIP = 2,
  SHAFFT_ERRSRC_HIPFFT = 3,
  SHAFFT_ERRSRC_FFTW = 4,
  SHAFFT_ERRSRC_SYSTEM = 5
} shafft_errsrc_t;

/* Opaque stand-in for MPI_Comm at one rank (0 = the calling process). */
typedef intptr_t shafft_comm_t;

int shafftSetStream(void* planPtr, void* stream);                     /* shafft.h:42  */
int shafftLastErrorStatus(void);                                      /* shafft.h:144 */
int shafftLastErrorSource(void);                                      /* shafft.h:151 */
int shafftLastErrorDomainCode(void);                                  /* shafft.h:166 */
int shafftLastErrorMessage(char* buf, int buflen);                    /* shafft.h:175 */
void shafftClearLastError(void);                                      /* shafft.h:183 */
const char* shafftErrorSourceName(int source);                        /* shafft.h:191 */
typedef enum { SHAFFT_MAXIMIZE_NDA, SHAFFT_MINIMIZE_NDA } shafft_decomposition_strategy_t;
/* One rank: the grid resolves to {1,...,1}, nda 0, subsize = size, offset 0,
 * commSize 1, whatever commDims/nda/strategy ask for (the reference's
 * commDims -> nda -> strategy fallback chain ends there at one rank,
 * src/cpp/configuration.cpp:530-600); memLimit is not applied at one rank
 * (:365, no Cartesian communicator).  ndim 1 needs commDims = {1}. */
int shafftConfigurationND(int ndim, int* size, shafft_t precision, int* commDims, int* nda,
                          size_t* subsize, size_t* offset, int* commSize,
                          shafft_decomposition_strategy_t strategy, size_t memLimit,
                          shafft_comm_t cComm);                       /* shafft.h:510 */
int shafftNDCreate(void** outPlan);                                   /* shafft.h:529 */
int shafftNDInit(void* planPtr, int ndim, int commDims[], int dimensions[], shafft_t precision,
                 shafft_comm_t cComm, shafft_transform_layout_t output_policy); /* :544 */
int shafftPlan(void* planPtr);                                        /* shafft.h:562 */
int shafftDestroy(void** planPtr);                                    /* shafft.h:573 */
int shafftGetLayout(void* planPtr, size_t* subsize, size_t* offset,
                    shafft_tensor_layout_t layout);                   /* shafft.h:585 */
int shafftGetAxes(void* planPtr, int* ca, int* da, shafft_tensor_layout_t layout); /* :597 */
int shafftGetAllocSize(void* plan, size_t* localAllocSize);           /* shafft.h:607 */
int shafftGetGlobalSize(void* plan, size_t* globalSize);              /* shafft.h:617 */
int shafftIsConfigured(void* plan, int* configured);                  /* shafft.h:627 */
int shafftIsActive(void* plan, int* active);                          /* shafft.h:637 */
int shafftSetBuffers(void* planPtr, void* data, void* work);          /* shafft.h:650 */
int shafftGetBuffers(void* planPtr, void** data, void** work);        /* shafft.h:664 */
int shafftExecute(void* planPtr, shafft_direction_t direction);       /* shafft.h:674 */
int shafftNormalize(void* planPtr);                                   /* shafft.h:683 */
int shafftConfiguration1D(size_t N, size_t* localN, size_t* localStart, size_t* localAllocSize,
                          shafft_t precision, shafft_comm_t cComm);   /* shafft.h:697 */
int shafft1DCreate(void** outPlan);                                   /* shafft.h:711 */
int shafft1DInit(void* planPtr, size_t N, size_t localN, size_t localStart, shafft_t precision,
                 shafft_comm_t cComm);                                /* shafft.h:725 */
int shafftAllocBufferF(size_t count, void** buf);                     /* shafft.h:738 */
int shafftAllocBufferD(size_t count, void** buf);                     /* shafft.h:750 */
int shafftFreeBufferF(void* buf);                                     /* shafft.h:758 */
int shafftFreeBufferD(void* buf);                                     /* shafft.h:766 */
int shafftCopyToBufferF(void* dst, const void* src, size_t count);    /* shafft.h:777 */
int shafftCopyToBufferD(void* dst, const void* src, size_t count);    /* shafft.h:788 */
int shafftCopyFromBufferF(void* dst, const void* src, size_t count);  /* shafft.h:799 */
int shafftCopyFromBufferD(void* dst, const void* src, size_t count);  /* shafft.h:810 */
const char* shafftGetBackendName(void);                               /* shafft.h:817 */
void shafftGetVersion(int* major, int* minor, int* patch);            /* shafft.h:827 */
const char* shafftGetVersionString(void);                             /* shafft.h:834 */
int shafftFinalize(void);                                             /* shafft.h:845 */

/* Configuration objects (shafft.h:203-475).  The structs keep the
 * reference's field names and order with MPI_Comm replaced by the one-rank
 * handle; arrays and strings in them are owned by the config (malloc'd by
 * Init, freed by Release).  At one rank Resolve gives commDims {1,...,1},
 * nda 0, commSize 1, both layouts = the whole tensor (offset 0, every axis
 * contiguous, none distributed), isActive 1, activeRank 0, activeSize 1,
 * nodeId 0, nodeCount 1; the hostname comes from gethostname and deviceName
 * from the current HIP device (NULL when there is none).  Flags as in
 * src/cpp/config_info.cpp:494-501: RESOLVED, plus CHANGED_COMM_DIMS /
 * CHANGED_NDA when the resolved grid or nda differs from the hints. */
typedef enum {
  SHAFFT_CONFIG_CHANGED_COMM_DIMS = (1 << 0),
  SHAFFT_CONFIG_CHANGED_NDA = (1 << 1),
  SHAFFT_CONFIG_INACTIVE_RANKS = (1 << 2),
  SHAFFT_CONFIG_RESOLVED = (1 << 3)
} shafft_config_flags_t;                                              /* shafft.h:208 */
typedef struct shafft_nd_layout_t {
  size_t* subsize;
  size_t* offset;
  size_t localElements;
  int hasLocalElements;
  int* contiguousAxes;  /* length ndim - nda */
  int* distributedAxes; /* length nda */
} shafft_nd_layout_t;                                                 /* shafft.h:224 */
typedef struct shafft_nd_config_t {
  size_t structSize; /* sizeof(shafft_nd_config_t) once initialised */
  int apiVersion;
  int ndim;
  shafft_t precision;
  size_t* globalShape;
  shafft_decomposition_strategy_t strategy;
  shafft_transform_layout_t outputPolicy;
  size_t memLimit;
  int* hintCommDims;
  int hintNda;
  int* commDims;
  int nda;
  int commSize;
  size_t allocElements;
  shafft_nd_layout_t initial;
  shafft_nd_layout_t output;
  int isActive;
  int activeRank;
  int activeSize;
  shafft_comm_t worldComm;
  shafft_comm_t activeComm;
  int nodeId;
  int nodeCount;
  char* hostname;
  size_t hostnameLen;
  char* deviceName;
  size_t deviceNameLen;
  int flags;
  int status;
  int reserved[8];
} shafft_nd_config_t;                                                 /* shafft.h:244 */
typedef struct shafft_1d_layout_t {
  size_t localSize;
  size_t localStart;
  int hasLocalElements;
} shafft_1d_layout_t;                                                 /* shafft.h:298 */
typedef struct shafft_1d_config_t {
  size_t structSize;
  int apiVersion;
  size_t globalSize;
  shafft_t precision;
  shafft_1d_layout_t initial;
  shafft_1d_layout_t output;
  size_t allocElements;
  int isActive;
  int activeRank;
  int activeSize;
  shafft_comm_t worldComm;
  shafft_comm_t activeComm;
  int nodeId;
  int nodeCount;
  char* hostname;
  size_t hostnameLen;
  char* deviceName;
  size_t deviceNameLen;
  int flags;
  int status;
  int reserved[8];
} shafft_1d_config_t;                                                 /* shafft.h:313 */
/* cfg must be zero-initialised; a second Init without Release is
 * ERR_INVALID_STATE; cComm must be 0 (ERR_INVALID_COMM). */
int shafftConfigNDInit(shafft_nd_config_t* cfg, int ndim, const size_t* globalShape,
                       shafft_t precision, const int* commDims, int hintNda,
                       shafft_decomposition_strategy_t strategy,
                       shafft_transform_layout_t outputPolicy, size_t memLimit,
                       shafft_comm_t cComm);                          /* shafft.h:372 */
void shafftConfigNDRelease(shafft_nd_config_t* cfg);                  /* shafft.h:394 */
int shafftConfigNDResolve(shafft_nd_config_t* cfg);                   /* shafft.h:406 */
int shafftNDInitFromConfig(void* planPtr, shafft_nd_config_t* cfg);   /* shafft.h:419 */
int shafftConfig1DInit(shafft_1d_config_t* cfg, size_t globalSize, shafft_t precision,
                       shafft_comm_t cComm);                          /* shafft.h:435 */
void shafftConfig1DRelease(shafft_1d_config_t* cfg);                  /* shafft.h:450 */
int shafftConfig1DResolve(shafft_1d_config_t* cfg);                   /* shafft.h:462 */
int shafft1DInitFromConfig(void* planPtr, shafft_1d_config_t* cfg);   /* shafft.h:475 */
/* The plan's communicator: the one-rank handle 0 (shafft.h:489). */
int shafftGetCommunicator(void* planPtr, shafft_comm_t* outComm);

/* Fortran-facing companions bound by the reference's Fortran module
 * (src/f03/shafftf03.h; src/f03/shafft.f03).  The Fortran communicator
 * handle is read through the pointer and must be 0 (the one-rank handle,
 * as cComm above); everything else forwards to the C entry point. */
typedef int shafft_fint_t; /* MPI_Fint */
int shafftConfigurationNDf03(int ndim, int* size, shafft_t precision, int* commDims, int* nda,
                             size_t* subsize, size_t* offset, int* commSize,
                             shafft_decomposition_strategy_t strategy, size_t memLimit,
                             shafft_fint_t* fHandle);                 /* shafftf03.h:11 */
int shafftNDInitf03(void* planPtr, int ndim, int commDims[], int dimensions[],
                    shafft_t precision, shafft_fint_t* fHandle,
                    shafft_transform_layout_t* output_policy);        /* shafftf03.h:23 */
int shafftConfiguration1Df03(size_t N, size_t* localN, size_t* localStart,
                             size_t* localAllocSize, shafft_t precision,
                             shafft_fint_t* fHandle);                 /* shafftf03.h:32 */
int shafft1DInitf03(void* planPtr, size_t N, size_t localN, size_t localStart,
                    shafft_t precision, shafft_fint_t* fHandle);      /* shafftf03.h:39 */
/* Config structs on the heap for the Fortran derived types (shafftf03.h:47-50). */
shafft_nd_config_t* shafftConfigNDAllocf03(void);
void shafftConfigNDFreef03(shafft_nd_config_t* cfg);
shafft_1d_config_t* shafftConfig1DAllocf03(void);
void shafftConfig1DFreef03(shafft_1d_config_t* cfg);
/* Config wrappers taking the Fortran handle (shafftf03.h:53-71). */
int shafftConfigNDInitf03(shafft_nd_config_t* cfg, int ndim, const size_t* globalShape,
                          shafft_t precision, const int* commDims, int hintNda,
                          shafft_decomposition_strategy_t strategy,
                          shafft_transform_layout_t outputPolicy, size_t memLimit,
                          shafft_fint_t* fComm);
int shafftConfigNDResolvef03(shafft_nd_config_t* cfg);
int shafftNDInitFromConfigf03(void* planPtr, shafft_nd_config_t* cfg);
int shafftConfig1DInitf03(shafft_1d_config_t* cfg, size_t globalSize, shafft_t precision,
                          shafft_fint_t* fComm);
int shafftConfig1DResolvef03(shafft_1d_config_t* cfg);
int shafft1DInitFromConfigf03(void* planPtr, shafft_1d_config_t* cfg);
int shafftGetCommunicatorf03(void* planPtr, shafft_fint_t* outFint);
/* Per-field accessors behind the Fortran derived-type sync
 * (shafftf03.h:76-146); a null config reads as 0 and writes nothing. */
int shafftConfigNDGetNdim(const shafft_nd_config_t* cfg);
int shafftConfigNDGetPrecision(const shafft_nd_config_t* cfg);
size_t shafftConfigNDGetAllocElements(const shafft_nd_config_t* cfg);
int shafftConfigNDGetIsActive(const shafft_nd_config_t* cfg);
int shafftConfigNDGetActiveRank(const shafft_nd_config_t* cfg);
int shafftConfigNDGetActiveSize(const shafft_nd_config_t* cfg);
int shafftConfigNDGetNda(const shafft_nd_config_t* cfg);
int shafftConfigNDGetCommSize(const shafft_nd_config_t* cfg);
int shafftConfigNDGetStatus(const shafft_nd_config_t* cfg);
int shafftConfigNDGetOutputPolicy(const shafft_nd_config_t* cfg);
int shafftConfigNDGetStrategy(const shafft_nd_config_t* cfg);
size_t shafftConfigNDGetMemLimit(const shafft_nd_config_t* cfg);
int shafftConfigNDGetHintNda(const shafft_nd_config_t* cfg);
int shafftConfigNDGetFlags(const shafft_nd_config_t* cfg);
shafft_fint_t shafftConfigNDGetWorldComm(const shafft_nd_config_t* cfg);
shafft_fint_t shafftConfigNDGetActiveComm(const shafft_nd_config_t* cfg);
int shafftConfigNDGetNodeId(const shafft_nd_config_t* cfg);
int shafftConfigNDGetNodeCount(const shafft_nd_config_t* cfg);
size_t shafftConfigNDGetHostnameLen(const shafft_nd_config_t* cfg);
void shafftConfigNDGetHostname(const shafft_nd_config_t* cfg, char* out, size_t maxLen);
size_t shafftConfigNDGetDeviceNameLen(const shafft_nd_config_t* cfg);
void shafftConfigNDGetDeviceName(const shafft_nd_config_t* cfg, char* out, size_t maxLen);
void shafftConfigNDGetGlobalShape(const shafft_nd_config_t* cfg, size_t* out);
void shafftConfigNDGetCommDims(const shafft_nd_config_t* cfg, int* out);
void shafftConfigNDGetHintCommDims(const shafft_nd_config_t* cfg, int* out);
void shafftConfigNDGetInitialSubsize(const shafft_nd_config_t* cfg, size_t* out);
void shafftConfigNDGetInitialOffset(const shafft_nd_config_t* cfg, size_t* out);
void shafftConfigNDGetOutputSubsize(const shafft_nd_config_t* cfg, size_t* out);
void shafftConfigNDGetOutputOffset(const shafft_nd_config_t* cfg, size_t* out);
void shafftConfigNDSetOutputPolicy(shafft_nd_config_t* cfg, int pol);
void shafftConfigNDSetStrategy(shafft_nd_config_t* cfg, int s);
void shafftConfigNDSetMemLimit(shafft_nd_config_t* cfg, size_t limit);
void shafftConfigNDSetHintNda(shafft_nd_config_t* cfg, int nda);
void shafftConfigNDSetHintCommDims(shafft_nd_config_t* cfg, const int* dims);
size_t shafftConfig1DGetGlobalSize(const shafft_1d_config_t* cfg);
int shafftConfig1DGetPrecision(const shafft_1d_config_t* cfg);
size_t shafftConfig1DGetAllocElements(const shafft_1d_config_t* cfg);
int shafftConfig1DGetIsActive(const shafft_1d_config_t* cfg);
int shafftConfig1DGetActiveRank(const shafft_1d_config_t* cfg);
int shafftConfig1DGetActiveSize(const shafft_1d_config_t* cfg);
int shafftConfig1DGetStatus(const shafft_1d_config_t* cfg);
int shafftConfig1DGetFlags(const shafft_1d_config_t* cfg);
shafft_fint_t shafftConfig1DGetWorldComm(const shafft_1d_config_t* cfg);
shafft_fint_t shafftConfig1DGetActiveComm(const shafft_1d_config_t* cfg);
int shafftConfig1DGetNodeId(const shafft_1d_config_t* cfg);
int shafftConfig1DGetNodeCount(const shafft_1d_config_t* cfg);
size_t shafftConfig1DGetHostnameLen(const shafft_1d_config_t* cfg);
void shafftConfig1DGetHostname(const shafft_1d_config_t* cfg, char* out, size_t maxLen);
size_t shafftConfig1DGetDeviceNameLen(const shafft_1d_config_t* cfg);
void shafftConfig1DGetDeviceName(const shafft_1d_config_t* cfg, char* out, size_t maxLen);
size_t shafftConfig1DGetInitialLocalSize(const shafft_1d_config_t* cfg);
size_t shafftConfig1DGetInitialLocalStart(const shafft_1d_config_t* cfg);
size_t shafftConfig1DGetOutputLocalSize(const shafft_1d_config_t* cfg);
size_t shafftConfig1DGetOutputLocalStart(const shafft_1d_config_t* cfg);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* SHAFFT_AMD_H */
