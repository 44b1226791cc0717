/* One-rank conformance run of SHAFFT's C API against libshafft_amd.so.
 *
 * Follows the call sequences of the reference's C integration tests
 * (tests/integration/c/test_c_api.c, test_c_roundtrip.c, test_c_FFT1D.c)
 * with the communicator argument 0: library info and error queries, plan
 * life cycle, invalid output policy, buffer helpers, buffer bookkeeping
 * across execute, layout / axes queries, configuration at one rank, N-D and
 * 1-D forward -> normalize -> backward -> normalize round trips, and the
 * closed forms of a delta and a plane wave (which fix the sign convention
 * and the 1/sqrt(N) normalisation).
 *
 * Prints "PASS <check>" / "FAIL <check> (<last error>)" per check and exits
 * with the number of failed checks.  Needs a GPU, except with the argument
 * "host" (info, errors, life cycle and configuration only).  Build:
 *   gcc -std=c99 -O1 -Iinclude tests/c/c_api_onerank.c \
 *       -Lshafft_amd -lshafft_amd -Wl,-rpath,<dir of the library> -lm
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "shafft_amd.h"

static const double kPi = 3.14159265358979323846;
static int g_fail = 0;

static void check(const char* name, int ok) {
  if (ok) {
    printf("PASS %s\n", name);
  } else {
    char msg[256] = {0};
    shafftLastErrorMessage(msg, (int)sizeof msg);
    printf("FAIL %s (%s)\n", name, msg);
    ++g_fail;
  }
  fflush(stdout);
}

/* Deterministic uniform[-1,1) values (64-bit LCG, top 53 bits). */
static uint64_t g_state = 42u;
static double next_uniform(void) {
  g_state = g_state * 6364136223846793005ull + 1442695040888963407ull;
  return 2.0 * (double)(g_state >> 11) * (1.0 / 9007199254740992.0) - 1.0;
}

/* ---------------------------------------------------------------- info */

static void test_info_and_errors(void) {
  const char* backend = shafftGetBackendName();
  check("backend name non-empty", backend != NULL && strlen(backend) > 0);
  int major = -1, minor = -1, patch = -1;
  shafftGetVersion(&major, &minor, &patch);
  check("version components non-negative", major >= 0 && minor >= 0 && patch >= 0);
  const char* ver = shafftGetVersionString();
  check("version string has a dot", ver != NULL && strchr(ver, '.') != NULL);

  shafftClearLastError();
  (void)shafftGetBackendName();
  check("no error after a successful call", shafftLastErrorStatus() == SHAFFT_SUCCESS);
  check("error source names",
        strlen(shafftErrorSourceName(SHAFFT_ERRSRC_NONE)) > 0 &&
            strlen(shafftErrorSourceName(SHAFFT_ERRSRC_MPI)) > 0 &&
            strlen(shafftErrorSourceName(SHAFFT_ERRSRC_HIP)) > 0);

  int rc = shafftNDInit(NULL, 3, NULL, NULL, SHAFFT_C2C, 0, SHAFFT_LAYOUT_REDISTRIBUTED);
  check("NDInit(NULL) fails with NULLPTR",
        rc == SHAFFT_ERR_NULLPTR && shafftLastErrorStatus() == SHAFFT_ERR_NULLPTR);
  char buf[256] = {0};
  check("error message recorded", shafftLastErrorMessage(buf, (int)sizeof buf) > 0);
  shafftClearLastError();
  check("error cleared", shafftLastErrorStatus() == SHAFFT_SUCCESS);
}

/* ----------------------------------------------------------- life cycle */

static void test_life_cycle(void) {
  void* plan = NULL;
  check("NDCreate", shafftNDCreate(&plan) == 0 && plan != NULL);
  check("Destroy nulls the handle", shafftDestroy(&plan) == 0 && plan == NULL);
  check("1DCreate", shafft1DCreate(&plan) == 0 && plan != NULL);
  check("Destroy 1-D", shafftDestroy(&plan) == 0 && plan == NULL);

  int dims[3] = {16, 16, 16}, comm_dims[3] = {1, 1, 1};
  shafftNDCreate(&plan);
  int rc = shafftNDInit(plan, 3, comm_dims, dims, SHAFFT_C2C, 0, (shafft_transform_layout_t)2);
  check("invalid output policy -> INVALID_LAYOUT", rc == SHAFFT_ERR_INVALID_LAYOUT);
  shafftDestroy(&plan);

  shafftNDCreate(&plan);
  rc = shafftPlan(plan);
  check("Plan before Init -> PLAN_NOT_INIT", rc == SHAFFT_ERR_PLAN_NOT_INIT);
  shafftDestroy(&plan);
}

/* -------------------------------------------------------------- buffers */

static void test_buffers(void) {
  const size_t n = 1000;
  void *df = NULL, *dd = NULL;
  check("AllocBufferF", shafftAllocBufferF(n, &df) == 0 && df != NULL);
  check("AllocBufferD", shafftAllocBufferD(n, &dd) == 0 && dd != NULL);
  float* hf = (float*)malloc(2 * n * sizeof(float));
  float* gf = (float*)calloc(2 * n, sizeof(float));
  double* hd = (double*)malloc(2 * n * sizeof(double));
  double* gd = (double*)calloc(2 * n, sizeof(double));
  for (size_t i = 0; i < 2 * n; ++i) {
    hf[i] = (float)next_uniform();
    hd[i] = next_uniform();
  }
  int rc = shafftCopyToBufferF(df, hf, n) | shafftCopyFromBufferF(gf, df, n);
  check("CopyTo/FromBufferF bitwise", rc == 0 && memcmp(hf, gf, 2 * n * sizeof(float)) == 0);
  rc = shafftCopyToBufferD(dd, hd, n) | shafftCopyFromBufferD(gd, dd, n);
  check("CopyTo/FromBufferD bitwise", rc == 0 && memcmp(hd, gd, 2 * n * sizeof(double)) == 0);
  check("FreeBufferF", shafftFreeBufferF(df) == 0);
  check("FreeBufferD", shafftFreeBufferD(dd) == 0);
  free(hf);
  free(gf);
  free(hd);
  free(gd);
}

/* Allocation edge cases of src/cpp/detail/buffer_utils.hpp:37-56 that make
 * no device call: count 0 is a null buffer and success; a byte count that
 * does not fit size_t is ERR_ALLOC. */
static void test_alloc_edges_host(void) {
  void* b = (void*)&g_fail;
  check("AllocBufferF(0) -> success, NULL", shafftAllocBufferF(0, &b) == 0 && b == NULL);
  b = (void*)&g_fail;
  check("AllocBufferD(SIZE_MAX) -> ALLOC, NULL",
        shafftAllocBufferD(SIZE_MAX, &b) == SHAFFT_ERR_ALLOC && b == NULL &&
            shafftLastErrorStatus() == SHAFFT_ERR_ALLOC);
  check("FreeBuffer(NULL) -> success", shafftFreeBufferF(NULL) == 0);
  float h[2] = {1.0f, 2.0f};
  check("CopyToBufferF count 0 -> success", shafftCopyToBufferF((void*)h, h, 0) == 0);
}

/* An allocation the device cannot hold fails with ERR_ALLOC and leaves the
 * library usable. */
static void test_alloc_too_large(void) {
  void* b = (void*)&g_fail;
  const size_t huge = (size_t)1 << 42; /* 64 TiB of complex128 */
  check("AllocBufferD(2^42) -> ALLOC, NULL", shafftAllocBufferD(huge, &b) == SHAFFT_ERR_ALLOC && b == NULL);
  void* ok = NULL;
  check("allocation after a failed one", shafftAllocBufferF(1024, &ok) == 0 && ok != NULL);
  shafftFreeBufferF(ok);
}

/* ------------------------------------------------------- configuration */

static void test_configuration(void) {
  int dims[3] = {64, 64, 32};
  for (int s = 0; s < 2; ++s) {
    int comm_dims[3] = {0, 0, 0}, nda = -1, comm_size = 0;
    size_t subsize[3] = {0, 0, 0}, offset[3] = {9, 9, 9};
    int rc = shafftConfigurationND(3, dims, SHAFFT_C2C, comm_dims, &nda, subsize, offset,
                                   &comm_size,
                                   s ? SHAFFT_MINIMIZE_NDA : SHAFFT_MAXIMIZE_NDA, 0, 0);
    check(s ? "ConfigurationND minimize at one rank" : "ConfigurationND maximize at one rank",
          rc == 0 && nda == 0 && comm_size == 1 && comm_dims[0] == 1 && comm_dims[1] == 1 &&
              comm_dims[2] == 1 && subsize[0] == 64 && subsize[1] == 64 && subsize[2] == 32 &&
              offset[0] == 0 && offset[1] == 0 && offset[2] == 0);
  }
  size_t ln = 0, ls = 9, la = 0;
  int rc = shafftConfiguration1D(1024, &ln, &ls, &la, SHAFFT_C2C, 0);
  check("Configuration1D at one rank", rc == 0 && ln == 1024 && ls == 0 && la >= 1024);
}

/* Configuration objects (shafft.h:203-475) at one rank: resolved fields,
 * flags, error cases, release. */
static void test_config_objects_host(void) {
  shafft_nd_config_t c;
  memset(&c, 0, sizeof c);
  const size_t shape[3] = {64, 48, 30};
  const int hint[3] = {0, 0, 0};
  int rc = shafftConfigNDInit(&c, 3, shape, SHAFFT_Z2Z, hint, 0, SHAFFT_MAXIMIZE_NDA,
                              SHAFFT_LAYOUT_REDISTRIBUTED, 0, 0);
  int ok = rc == 0 && c.structSize == sizeof c && c.ndim == 3 && c.nda == 0 && c.commSize == 1 &&
           c.allocElements == 64 * 48 * 30 && c.isActive == 1 && c.activeRank == 0 &&
           c.activeSize == 1 && c.nodeCount == 1 && c.hostname != NULL &&
           c.hostnameLen == strlen(c.hostname) && (c.flags & SHAFFT_CONFIG_RESOLVED) &&
           (c.flags & SHAFFT_CONFIG_CHANGED_COMM_DIMS) && !(c.flags & SHAFFT_CONFIG_CHANGED_NDA) &&
           c.status == 0;
  for (int i = 0; ok && i < 3; ++i)
    ok = c.commDims[i] == 1 && c.initial.subsize[i] == shape[i] && c.initial.offset[i] == 0 &&
         c.output.subsize[i] == shape[i] && c.initial.contiguousAxes[i] == i &&
         c.globalShape[i] == shape[i];
  ok = ok && c.initial.localElements == c.allocElements && c.initial.hasLocalElements == 1;
  check("ConfigNDInit resolves the one-rank layout", ok);
  check("ConfigNDInit twice -> INVALID_STATE",
        shafftConfigNDInit(&c, 3, shape, SHAFFT_Z2Z, hint, 0, SHAFFT_MAXIMIZE_NDA,
                           SHAFFT_LAYOUT_REDISTRIBUTED, 0, 0) == SHAFFT_ERR_INVALID_STATE);
  check("ConfigNDResolve again", shafftConfigNDResolve(&c) == 0 && (c.flags & SHAFFT_CONFIG_RESOLVED));
  shafftConfigNDRelease(&c);
  check("ConfigNDRelease zero-fills", c.structSize == 0 && c.globalShape == NULL && c.hostname == NULL);
  shafftConfigNDRelease(&c); /* idempotent */
  check("ConfigNDResolve after release -> INVALID_STATE",
        shafftConfigNDResolve(&c) == SHAFFT_ERR_INVALID_STATE);

  const int ones[3] = {1, 1, 1};
  memset(&c, 0, sizeof c);
  rc = shafftConfigNDInit(&c, 3, shape, SHAFFT_C2C, ones, 0, SHAFFT_MINIMIZE_NDA,
                          SHAFFT_LAYOUT_INITIAL, 0, 0);
  check("ConfigNDInit with the grid of ones: no CHANGED flags",
        rc == 0 && c.flags == SHAFFT_CONFIG_RESOLVED && c.outputPolicy == SHAFFT_LAYOUT_INITIAL);
  shafftConfigNDRelease(&c);

  const size_t bad[2] = {16, 0};
  memset(&c, 0, sizeof c);
  check("ConfigNDInit zero extent -> INVALID_DIM, released",
        shafftConfigNDInit(&c, 2, bad, SHAFFT_C2C, NULL, 0, SHAFFT_MAXIMIZE_NDA,
                           SHAFFT_LAYOUT_REDISTRIBUTED, 0, 0) == SHAFFT_ERR_INVALID_DIM &&
            c.structSize == 0);
  check("ConfigNDInit comm != 0 -> INVALID_COMM",
        shafftConfigNDInit(&c, 3, shape, SHAFFT_C2C, NULL, 0, SHAFFT_MAXIMIZE_NDA,
                           SHAFFT_LAYOUT_REDISTRIBUTED, 0, 7) == SHAFFT_ERR_INVALID_COMM);

  shafft_1d_config_t d;
  memset(&d, 0, sizeof d);
  rc = shafftConfig1DInit(&d, 1000003, SHAFFT_C2C, 0);
  check("Config1DInit resolves the one-rank layout",
        rc == 0 && d.structSize == sizeof d && d.globalSize == 1000003 &&
            d.initial.localSize == 1000003 && d.initial.localStart == 0 &&
            d.output.localSize == 1000003 && d.allocElements == 1000003 && d.isActive == 1 &&
            d.flags == SHAFFT_CONFIG_RESOLVED);
  shafftConfig1DRelease(&d);
  check("Config1DRelease zero-fills", d.structSize == 0 && d.hostname == NULL);
  check("Config1DInit size 0 -> INVALID_DIM", shafftConfig1DInit(&d, 0, SHAFFT_C2C, 0) == SHAFFT_ERR_INVALID_DIM);

  void* plan = NULL;
  shafft_comm_t comm = 99;
  shafftNDCreate(&plan);
  check("GetCommunicator before init -> PLAN_NOT_INIT",
        shafftGetCommunicator(plan, &comm) == SHAFFT_ERR_PLAN_NOT_INIT);
  memset(&c, 0, sizeof c);
  shafftConfigNDInit(&c, 3, shape, SHAFFT_C2C, NULL, 0, SHAFFT_MAXIMIZE_NDA,
                     SHAFFT_LAYOUT_REDISTRIBUTED, 0, 0);
  size_t alloc = 0, g = 0;
  rc = shafftNDInitFromConfig(plan, &c);
  check("NDInitFromConfig configures the plan",
        rc == 0 && shafftGetAllocSize(plan, &alloc) == 0 && alloc == c.allocElements &&
            shafftGetGlobalSize(plan, &g) == 0 && g == c.allocElements);
  check("GetCommunicator -> the one-rank handle", shafftGetCommunicator(plan, &comm) == 0 && comm == 0);
  shafftConfigNDRelease(&c);
  shafftDestroy(&plan);
}

/* Plans made from config objects run the same round trips. */
static void test_config_objects_gpu(void) {
  shafft_nd_config_t c;
  memset(&c, 0, sizeof c);
  const size_t shape[3] = {24, 40, 18};
  void *plan = NULL, *data = NULL, *work = NULL, *d = NULL, *w = NULL;
  int rc = shafftConfigNDInit(&c, 3, shape, SHAFFT_Z2Z, NULL, 0, SHAFFT_MAXIMIZE_NDA,
                              SHAFFT_LAYOUT_REDISTRIBUTED, 0, 0);
  rc |= shafftNDCreate(&plan) | shafftNDInitFromConfig(plan, &c);
  rc |= shafftAllocBufferD(c.allocElements, &data) | shafftAllocBufferD(c.allocElements, &work);
  rc |= shafftSetBuffers(plan, data, work) | shafftPlan(plan);
  const size_t n = c.allocElements;
  double* x = (double*)malloc(2 * n * sizeof(double));
  double* y = (double*)malloc(2 * n * sizeof(double));
  for (size_t i = 0; i < 2 * n; ++i) x[i] = next_uniform();
  rc |= shafftCopyToBufferD(data, x, n);
  rc |= shafftExecute(plan, SHAFFT_FORWARD) | shafftNormalize(plan);
  rc |= shafftExecute(plan, SHAFFT_BACKWARD) | shafftNormalize(plan);
  rc |= shafftGetBuffers(plan, &d, &w) | shafftCopyFromBufferD(y, d, n);
  double num = 0, den = 0;
  for (size_t i = 0; i < 2 * n; ++i) {
    num += (y[i] - x[i]) * (y[i] - x[i]);
    den += x[i] * x[i];
  }
  printf("  config-object N-D round trip rel L2 %.3e\n", sqrt(num / den));
  check("N-D round trip of a plan from a config object", rc == 0 && sqrt(num / den) <= 1e-12);
  free(x);
  free(y);
  shafftFreeBufferD(data);
  shafftFreeBufferD(work);
  shafftDestroy(&plan);
  shafftConfigNDRelease(&c);
}

/* ------------------------------------------------------ N-D transforms */

typedef struct {
  void* plan;
  void *data, *work;
  size_t alloc;
} nd_case;

static int nd_setup(nd_case* c, int ndim, int* dims, shafft_t prec) {
  int comm_dims[8] = {1, 1, 1, 1, 1, 1, 1, 1};
  memset(c, 0, sizeof *c);
  if (shafftNDCreate(&c->plan)) return 1;
  if (shafftNDInit(c->plan, ndim, comm_dims, dims, prec, 0, SHAFFT_LAYOUT_REDISTRIBUTED)) return 2;
  if (shafftGetAllocSize(c->plan, &c->alloc) || c->alloc == 0) return 3;
  int rc = prec == SHAFFT_C2C ? shafftAllocBufferF(c->alloc, &c->data) |
                                    shafftAllocBufferF(c->alloc, &c->work)
                              : shafftAllocBufferD(c->alloc, &c->data) |
                                    shafftAllocBufferD(c->alloc, &c->work);
  if (rc) return 4;
  /* buffers before plan(), as the reference's C tests do */
  if (shafftSetBuffers(c->plan, c->data, c->work)) return 5;
  if (shafftPlan(c->plan)) return 6;
  return 0;
}

static void nd_teardown(nd_case* c, shafft_t prec) {
  if (prec == SHAFFT_C2C) {
    if (c->data) shafftFreeBufferF(c->data);
    if (c->work) shafftFreeBufferF(c->work);
  } else {
    if (c->data) shafftFreeBufferD(c->data);
    if (c->work) shafftFreeBufferD(c->work);
  }
  shafftDestroy(&c->plan);
}

/* Upload n complex values held as doubles in the plan's precision. */
static int upload(void* dst, const double* h, size_t n, shafft_t prec) {
  if (prec == SHAFFT_Z2Z) return shafftCopyToBufferD(dst, h, n);
  float* f = (float*)malloc(2 * n * sizeof(float));
  for (size_t i = 0; i < 2 * n; ++i) f[i] = (float)h[i];
  int rc = shafftCopyToBufferF(dst, f, n);
  free(f);
  return rc;
}

static int download(double* h, const void* src, size_t n, shafft_t prec) {
  if (prec == SHAFFT_Z2Z) return shafftCopyFromBufferD(h, src, n);
  float* f = (float*)malloc(2 * n * sizeof(float));
  int rc = shafftCopyFromBufferF(f, src, n);
  for (size_t i = 0; i < 2 * n; ++i) h[i] = f[i];
  free(f);
  return rc;
}

static double rel_l2(const double* a, const double* b, size_t n2) {
  double num = 0, den = 0;
  for (size_t i = 0; i < n2; ++i) {
    num += (a[i] - b[i]) * (a[i] - b[i]);
    den += b[i] * b[i];
  }
  return sqrt(num / (den > 0 ? den : 1));
}

static void test_nd_roundtrip(int ndim, int* dims, shafft_t prec, const char* name) {
  nd_case c;
  int rc = nd_setup(&c, ndim, dims, prec);
  if (rc) {
    check(name, 0);
    nd_teardown(&c, prec);
    return;
  }
  size_t n = c.alloc;
  double* x = (double*)malloc(2 * n * sizeof(double));
  double* y = (double*)malloc(2 * n * sizeof(double));
  for (size_t i = 0; i < 2 * n; ++i) x[i] = next_uniform();
  if (prec == SHAFFT_C2C)
    for (size_t i = 0; i < 2 * n; ++i) x[i] = (float)x[i];
  void *d = NULL, *w = NULL;
  rc = upload(c.data, x, n, prec);
  rc |= shafftExecute(c.plan, SHAFFT_FORWARD) | shafftNormalize(c.plan);
  rc |= shafftExecute(c.plan, SHAFFT_BACKWARD) | shafftNormalize(c.plan);
  rc |= shafftGetBuffers(c.plan, &d, &w);
  rc |= download(y, d, n, prec);
  double err = rel_l2(y, x, 2 * n);
  double tol = prec == SHAFFT_C2C ? 1e-5 : 1e-12;
  printf("  %s: round-trip rel L2 %.3e (tol %.0e)\n", name, err, tol);
  check(name, rc == 0 && err <= tol);
  free(x);
  free(y);
  nd_teardown(&c, prec);
}

/* x[n] = exp(+2 pi i k0.n / N) -> forward + normalize = sqrt(N) at k0 only. */
static void test_nd_plane_wave(void) {
  int dims[3] = {12, 10, 14};
  const int k0[3] = {5, 3, 11};
  nd_case c;
  if (nd_setup(&c, 3, dims, SHAFFT_Z2Z)) {
    check("N-D plane wave", 0);
    nd_teardown(&c, SHAFFT_Z2Z);
    return;
  }
  size_t n = c.alloc;
  double* x = (double*)malloc(2 * n * sizeof(double));
  size_t i = 0;
  for (int a = 0; a < dims[0]; ++a)
    for (int b = 0; b < dims[1]; ++b)
      for (int e = 0; e < dims[2]; ++e, ++i) {
        /* exact phase in turns: (k0.n mod N_axis)/N_axis summed per axis */
        double t = (double)((k0[0] * a) % dims[0]) / dims[0] +
                   (double)((k0[1] * b) % dims[1]) / dims[1] +
                   (double)((k0[2] * e) % dims[2]) / dims[2];
        x[2 * i] = cos(2 * kPi * t);
        x[2 * i + 1] = sin(2 * kPi * t);
      }
  void *d = NULL, *w = NULL;
  int rc = upload(c.data, x, n, SHAFFT_Z2Z);
  rc |= shafftExecute(c.plan, SHAFFT_FORWARD) | shafftNormalize(c.plan);
  rc |= shafftGetBuffers(c.plan, &d, &w);
  rc |= download(x, d, n, SHAFFT_Z2Z);
  size_t peak = ((size_t)k0[0] * dims[1] + k0[1]) * dims[2] + k0[2];
  double sq = sqrt((double)n), worst = 0;
  for (i = 0; i < n; ++i) {
    double re = x[2 * i] - (i == peak ? sq : 0.0), im = x[2 * i + 1];
    double e = sqrt(re * re + im * im);
    if (e > worst) worst = e;
  }
  printf("  N-D plane wave: max |error| %.3e (peak %.6f)\n", worst, sq);
  check("N-D plane wave (sign and 1/sqrt(N))", rc == 0 && worst <= 1e-12 * sq);
  free(x);
  nd_teardown(&c, SHAFFT_Z2Z);
}

static void test_nd_queries(void) {
  int dims[3] = {8, 6, 10};
  nd_case c;
  int ok = nd_setup(&c, 3, dims, SHAFFT_C2C) == 0;
  size_t g = 0, sub[3] = {0, 0, 0}, off[3] = {9, 9, 9};
  int ca[3] = {-1, -1, -1}, da[3] = {-1, -1, -1}, conf = 0, act = 0;
  ok = ok && shafftGetGlobalSize(c.plan, &g) == 0 && g == 480 && c.alloc == 480;
  ok = ok && shafftGetLayout(c.plan, sub, off, SHAFFT_TENSOR_LAYOUT_INITIAL) == 0 &&
       sub[0] == 8 && sub[1] == 6 && sub[2] == 10 && off[0] == 0 && off[1] == 0 && off[2] == 0;
  ok = ok && shafftGetAxes(c.plan, ca, da, SHAFFT_TENSOR_LAYOUT_CURRENT) == 0 && ca[0] == 0 &&
       ca[1] == 1 && ca[2] == 2;
  ok = ok && shafftIsConfigured(c.plan, &conf) == 0 && conf == 1;
  ok = ok && shafftIsActive(c.plan, &act) == 0 && act == 1;
  check("N-D size / layout / axes / state queries", ok);

  void *d0 = NULL, *w0 = NULL, *d1 = NULL, *w1 = NULL;
  ok = shafftGetBuffers(c.plan, &d0, &w0) == 0 && d0 == c.data && w0 == c.work;
  check("GetBuffers returns the buffers set", ok);
  ok = shafftExecute(c.plan, SHAFFT_FORWARD) == 0 && shafftGetBuffers(c.plan, &d1, &w1) == 0 &&
       ((d1 == c.data && w1 == c.work) || (d1 == c.work && w1 == c.data));
  check("GetBuffers after execute is the same pair", ok);
  /* the pair is owned by this test; free it by the original names */
  nd_teardown(&c, SHAFFT_C2C);
}

/* ------------------------------------------------------ 1-D transforms */

static void test_1d(size_t N, shafft_t prec, int delta, const char* name) {
  size_t ln = 0, ls = 0, la = 0, alloc = 0, gn = 0, pln = 0, pls = 9;
  void *plan = NULL, *data = NULL, *work = NULL, *d = NULL, *w = NULL;
  int rc = shafftConfiguration1D(N, &ln, &ls, &la, prec, 0);
  rc |= shafft1DCreate(&plan);
  rc |= shafft1DInit(plan, N, ln, ls, prec, 0);
  rc |= shafftGetAllocSize(plan, &alloc);
  rc |= shafftGetGlobalSize(plan, &gn);
  rc |= shafftGetLayout(plan, &pln, &pls, SHAFFT_TENSOR_LAYOUT_INITIAL);
  if (rc || alloc < N || gn != N || pln != ln || pls != ls) {
    check(name, 0);
    shafftDestroy(&plan);
    return;
  }
  if (prec == SHAFFT_C2C)
    rc = shafftAllocBufferF(alloc, &data) | shafftAllocBufferF(alloc, &work);
  else
    rc = shafftAllocBufferD(alloc, &data) | shafftAllocBufferD(alloc, &work);
  rc |= shafftSetBuffers(plan, data, work) | shafftPlan(plan);
  double* x = (double*)calloc(2 * alloc, sizeof(double));
  double* y = (double*)calloc(2 * alloc, sizeof(double));
  if (delta) {
    x[0] = 1.0;
  } else {
    for (size_t i = 0; i < 2 * N; ++i) x[i] = next_uniform();
    if (prec == SHAFFT_C2C)
      for (size_t i = 0; i < 2 * N; ++i) x[i] = (float)x[i];
  }
  rc |= upload(data, x, alloc, prec);
  rc |= shafftExecute(plan, SHAFFT_FORWARD) | shafftNormalize(plan);
  double tol = prec == SHAFFT_C2C ? 1e-5 : 1e-12;
  double err;
  if (delta) {
    /* delta -> constant 1/sqrt(N) */
    rc |= shafftGetBuffers(plan, &d, &w) | download(y, d, N, prec);
    double c = 1.0 / sqrt((double)N), worst = 0;
    for (size_t i = 0; i < N; ++i) {
      double e = hypot(y[2 * i] - c, y[2 * i + 1]);
      if (e > worst) worst = e;
    }
    err = worst / c;
  } else {
    rc |= shafftExecute(plan, SHAFFT_BACKWARD) | shafftNormalize(plan);
    rc |= shafftGetBuffers(plan, &d, &w) | download(y, d, N, prec);
    err = rel_l2(y, x, 2 * N);
  }
  printf("  %s: error %.3e (tol %.0e)\n", name, err, tol);
  check(name, rc == 0 && err <= tol);
  free(x);
  free(y);
  if (prec == SHAFFT_C2C) {
    shafftFreeBufferF(data);
    shafftFreeBufferF(work);
  } else {
    shafftFreeBufferD(data);
    shafftFreeBufferD(work);
  }
  shafftDestroy(&plan);
}

int main(int argc, char** argv) {
  /* "host": only the checks that make no device call (no GPU needed) */
  const int host_only = argc > 1 && strcmp(argv[1], "host") == 0;
  test_info_and_errors();
  test_life_cycle();
  test_configuration();
  test_alloc_edges_host();
  test_config_objects_host();
  if (host_only) {
    printf("%d check(s) failed\n", g_fail);
    return g_fail;
  }
  test_buffers();
  test_alloc_too_large();
  test_config_objects_gpu();
  {
    int d1[3] = {32, 32, 32};
    test_nd_roundtrip(3, d1, SHAFFT_C2C, "N-D round trip 32x32x32 complex64");
    int d2[3] = {64, 64, 32};
    test_nd_roundtrip(3, d2, SHAFFT_Z2Z, "N-D round trip 64x64x32 complex128");
    int d3[2] = {96, 210};
    test_nd_roundtrip(2, d3, SHAFFT_C2C, "N-D round trip 96x210 complex64");
    int d4[4] = {6, 5, 7, 9};
    test_nd_roundtrip(4, d4, SHAFFT_Z2Z, "N-D round trip 6x5x7x9 complex128");
  }
  test_nd_plane_wave();
  test_nd_queries();
  test_1d(1024, SHAFFT_C2C, 0, "1-D round trip N=1024 complex64");
  test_1d(2048, SHAFFT_Z2Z, 0, "1-D round trip N=2048 complex128");
  test_1d(1009, SHAFFT_Z2Z, 0, "1-D round trip N=1009 (prime) complex128");
  test_1d(512, SHAFFT_C2C, 1, "1-D delta N=512 complex64");
  shafftFinalize();
  printf("%d check(s) failed\n", g_fail);
  return g_fail;
}
