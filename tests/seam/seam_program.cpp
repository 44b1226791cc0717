// Conformance program for the gfx950 seam adapter
// (shafft_amd/csrc/seam/gfx950_method): calls fftndPlan / fftndSetStream /
// fftndExecute / fftndDestroy, normalizeComplex*, and the shafft::fft1d*
// functions exactly as SHAFFT's core does (src/cpp/shafft_internal.cpp:
// 40,94,149-172,205,351-374,567-638,670-783): handles from
// `new FFTNDHandle[n]()`, the N-D handle passed by value to execute, data and
// work passed as void*& and swapped behind the caller's back.
//
// Checks, each printed PASS or FAIL:
//   * the seam-level partial-axis DFTs of tests/unit/test_fft_partial.cpp:
//     231-237 (input sin(0.01 i) + i cos(0.02 i), forward and backward,
//     1e-10 absolute against a direct DFT in long double);
//   * the buffer contract: the result behind `data`, {data, work} a
//     permutation of the caller's two buffers;
//   * streams as tests/integration/cpp/test_backend.cpp:251-450: a created
//     stream, switching between two streams and back to the default stream
//     all give bit-identical results;
//   * normalizeComplexFloat/Double scale exactly (x * f);
//   * FFT1D layout, plan, execute against the direct DFT, normalize, round
//     trip, destroy, and the errors for a destroyed handle and a null
//     communicator.
// Argument "multirank" (under mpiexec -n 2): each rank runs the N-D seam
// cases on its own (the N-D seam is local; the core exchanges slabs between
// stages), and FFT1D on the two-rank communicator is refused with
// ERR_NOT_IMPL (this backend runs the 1-D transform on one GPU).
#include <mpi.h>
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <vector>

#include "common/normalize.hpp"
#include "fft1d_method.hpp"
#include "fftnd_method.hpp"

namespace {

int failures = 0;

void report(bool ok, const char* what) {
  std::printf("%s %s\n", ok ? "PASS" : "FAIL", what);
  if (!ok) ++failures;
}

using cd = std::complex<double>;

// Direct DFT over consecutive axes [a0, a1] of a row-major tensor.
std::vector<cd> dft_axes(const std::vector<cd>& x, const std::vector<int>& dims, int a0, int a1,
                         int sign) {
  std::vector<cd> cur = x;
  for (int ax = a0; ax <= a1; ++ax) {
    long long outer = 1, inner = 1;
    for (int i = 0; i < ax; ++i) outer *= dims[i];
    for (size_t i = ax + 1; i < dims.size(); ++i) inner *= dims[i];
    const int n = dims[ax];
    std::vector<cd> nxt(cur.size());
    for (long long o = 0; o < outer; ++o)
      for (long long s = 0; s < inner; ++s)
        for (int k = 0; k < n; ++k) {
          std::complex<long double> acc = 0;
          for (int j = 0; j < n; ++j) {
            const long double a = sign * 2.0L * 3.141592653589793238462643383279502884L *
                                  (long double)((long long)j * k % n) / n;
            const cd v = cur[(o * n + j) * inner + s];
            acc += std::complex<long double>(v.real(), v.imag()) *
                   std::complex<long double>(cosl(a), sinl(a));
          }
          nxt[(o * n + k) * inner + s] = cd((double)acc.real(), (double)acc.imag());
        }
    cur.swap(nxt);
  }
  return cur;
}

std::vector<cd> partial_input(size_t n) {
  std::vector<cd> x(n);
  for (size_t i = 0; i < n; ++i) x[i] = cd(std::sin(0.01 * i), std::cos(0.02 * i));
  return x;
}

struct DevBuf {
  void* p = nullptr;
  size_t n = 0;
  explicit DevBuf(size_t count) : n(count) { shafftAllocBufferD(count, &p); }
  ~DevBuf() { shafftFreeBufferD(p); }
  void put(const std::vector<cd>& v) { shafftCopyToBufferD(p, v.data(), v.size()); }
};

std::vector<cd> get(const void* p, size_t n) {
  std::vector<cd> v(n);
  shafftCopyFromBufferD(v.data(), p, n);
  return v;
}

double max_abs_diff(const std::vector<cd>& a, const std::vector<cd>& b) {
  double m = 0;
  for (size_t i = 0; i < a.size(); ++i) m = std::max(m, std::abs(a[i] - b[i]));
  return m;
}

void nd_partial_cases() {
  struct Case {
    std::vector<int> dims;
    int a0, a1;
  };
  const std::vector<Case> cases = {{{2, 3, 4, 5}, 0, 2}, {{4, 3, 2, 3, 2}, 1, 3}, {{2, 3, 4, 5}, 2, 3},
                                   {{2, 3, 4, 2}, 0, 3}, {{3, 2, 2, 2, 2}, 1, 4}};
  FFTNDHandle* handles = new FFTNDHandle[cases.size()]();
  bool fresh = true;
  for (size_t c = 0; c < cases.size(); ++c)
    fresh &= handles[c].plan == nullptr && handles[c].stream == nullptr &&
             handles[c].comm == MPI_COMM_NULL;
  report(fresh, "new FFTNDHandle[n]() is value-initialised (plan, stream, comm)");
  bool all_ok = true, swap_ok = true;
  for (size_t c = 0; c < cases.size(); ++c) {
    std::vector<int> dims = cases[c].dims;
    std::vector<int> ta;
    for (int a = cases[c].a0; a <= cases[c].a1; ++a) ta.push_back(a);
    size_t n = 1;
    for (int d : dims) n *= d;
    if (fftndPlan(handles[c], (int)ta.size(), ta.data(), (int)dims.size(), dims.data(),
                  shafft::FFTType::Z2Z, nullptr, nullptr) != 0) {
      all_ok = false;
      continue;
    }
    const std::vector<cd> x = partial_input(n);
    for (int dir = 0; dir < 2; ++dir) {
      DevBuf a(n), b(n);
      a.put(x);
      void* data = a.p;
      void* work = b.p;
      const int rc = fftndExecute(handles[c], data, work,
                                  dir ? shafft::FFTDirection::BACKWARD : shafft::FFTDirection::FORWARD);
      swap_ok &= rc == 0 && ((data == a.p && work == b.p) || (data == b.p && work == a.p));
      const double err = max_abs_diff(get(data, n), dft_axes(x, dims, cases[c].a0, cases[c].a1,
                                                             dir ? +1 : -1));
      std::printf("  partial case %zu dir %d: max abs err %.2e\n", c, dir, err);
      all_ok &= rc == 0 && err < 1e-10;
    }
  }
  report(all_ok, "partial-axis seam DFTs within 1e-10 (test_fft_partial.cpp:231-237)");
  report(swap_ok, "execute leaves the result behind data, {data, work} swapped in place");
  bool destroyed = true;
  for (size_t c = 0; c < cases.size(); ++c) destroyed &= fftndDestroy(handles[c]) == 0;
  delete[] handles;
  report(destroyed, "fftndDestroy of every subplan");

  FFTNDHandle bad;
  int d2[2] = {4, 0}, t2[2] = {0, 1};
  report(fftndPlan(bad, 2, t2, 2, d2, shafft::FFTType::C2C, nullptr, nullptr) ==
             SHAFFT_ERR_INVALID_DIM, "zero extent -> ERR_INVALID_DIM");
  int d3[3] = {4, 4, 4}, t3[2] = {0, 2};
  report(fftndPlan(bad, 2, t3, 3, d3, shafft::FFTType::C2C, nullptr, nullptr) ==
             SHAFFT_ERR_INVALID_DIM, "non-consecutive axes -> ERR_INVALID_DIM");
}

void streams() {
  std::vector<int> dims = {16, 12, 10};
  int ta[3] = {0, 1, 2};
  const size_t n = 16 * 12 * 10;
  FFTNDHandle h;
  const bool planned = fftndPlan(h, 3, ta, 3, dims.data(), shafft::FFTType::Z2Z, nullptr, nullptr) == 0;
  const std::vector<cd> x = partial_input(n);
  auto run = [&](hipStream_t s) {
    fftndSetStream(h, s);
    DevBuf a(n), b(n);
    a.put(x);
    void* data = a.p;
    void* work = b.p;
    fftndExecute(h, data, work, shafft::FFTDirection::FORWARD);
    return get(data, n);
  };
  hipStream_t s1 = nullptr, s2 = nullptr;
  const bool created = hipStreamCreateWithFlags(&s1, hipStreamNonBlocking) == hipSuccess &&
                       hipStreamCreateWithFlags(&s2, hipStreamNonBlocking) == hipSuccess;
  report(created, "two non-blocking streams created");
  const std::vector<cd> r0 = run(nullptr);
  const std::vector<cd> r1 = run(s1);
  const std::vector<cd> r2 = run(s2);
  const std::vector<cd> r3 = run(s1);
  const std::vector<cd> r4 = run(nullptr);
  report(planned && h.stream == nullptr, "handle stream back to the default stream");
  report(max_abs_diff(r0, dft_axes(x, dims, 0, 2, -1)) < 1e-10, "default-stream result vs direct DFT");
  const bool same = std::memcmp(r0.data(), r1.data(), n * sizeof(cd)) == 0 &&
                    std::memcmp(r0.data(), r2.data(), n * sizeof(cd)) == 0 &&
                    std::memcmp(r0.data(), r3.data(), n * sizeof(cd)) == 0 &&
                    std::memcmp(r0.data(), r4.data(), n * sizeof(cd)) == 0;
  report(same, "created stream, stream switching and default stream: identical bits");

  // normalize on a stream (asynchronous): x * f exactly
  DevBuf a(n);
  a.put(x);
  const double f = 1.0 / std::sqrt(1920.0);
  int rc = normalizeComplexDouble(f, n, a.p, s1);
  if (hipStreamSynchronize(s1) != hipSuccess) rc = -1;
  const std::vector<cd> y = get(a.p, n);
  bool exact = rc == 0;
  for (size_t i = 0; i < n; ++i) exact &= y[i] == cd(x[i].real() * f, x[i].imag() * f);
  report(exact, "normalizeComplexDouble on a stream: x * f bit for bit");
  void* pf = nullptr;
  shafftAllocBufferF(4, &pf);
  const float xf[8] = {1, 2, 3, 4, 5, 6, 7, 8};
  shafftCopyToBufferF(pf, xf, 4);
  const int rcf = normalizeComplexFloat(0.5f, 4, pf);
  float yf[8];
  shafftCopyFromBufferF(yf, pf, 4);
  bool fexact = rcf == 0;
  for (int i = 0; i < 8; ++i) fexact &= yf[i] == xf[i] * 0.5f;
  report(fexact, "normalizeComplexFloat on the default stream");
  shafftFreeBufferF(pf);
  fftndDestroy(h);
  report(hipStreamDestroy(s1) == hipSuccess && hipStreamDestroy(s2) == hipSuccess, "streams destroyed");
}

void fft1d() {
  const size_t N = 1000;
  size_t ln = 0, ls = 9, la = 0, lnt = 0, lst = 9;
  const int q = shafft::fft1dQueryLayout(N, ln, ls, la, lnt, lst, MPI_COMM_WORLD);
  report(q == 0 && ln == N && ls == 0 && la >= N && lnt == N && lst == 0,
         "fft1dQueryLayout at one rank: localN = N, localStart = 0");
  shafft::FFT1DHandle* h = new shafft::FFT1DHandle();
  const int p = shafft::fft1dPlan(*h, N, shafft::FFTType::Z2Z, MPI_COMM_WORLD, nullptr, nullptr);
  report(p == 0 && h->internal && h->globalN == N && h->localNInit == N && h->localAllocSize >= N,
         "fft1dPlan fills the handle");
  hipStream_t s = nullptr;
  const bool created = hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess;
  report(created && shafft::fft1dSetStream(*h, s) == 0 && h->stream == s, "fft1dSetStream");
  std::vector<cd> x = partial_input(N);
  DevBuf a(la), b(la);
  a.put(x);
  int rc = shafft::fft1dExecute(*h, a.p, b.p, shafft::FFTDirection::FORWARD);
  const double e1 = max_abs_diff(get(b.p, N), dft_axes(x, {(int)N}, 0, 0, -1));
  report(rc == 0 && e1 < 1e-10, "fft1dExecute forward vs direct DFT (1e-10)");
  report(get(a.p, N) == x, "fft1dExecute keeps its input");
  rc = shafft::fft1dNormalize(*h, b.p, 1, N);
  rc |= shafft::fft1dExecute(*h, b.p, a.p, shafft::FFTDirection::BACKWARD);
  rc |= shafft::fft1dNormalize(*h, a.p, 1, N);
  report(rc == 0 && max_abs_diff(get(a.p, N), x) < 1e-12, "FFT1D forward/normalize/backward/normalize");
  report(shafft::fft1dSynchronize(*h) == 0, "fft1dSynchronize");
  report(shafft::fft1dDestroy(*h) == 0 && h->internal == nullptr, "fft1dDestroy");
  report(shafft::fft1dExecute(*h, a.p, b.p, shafft::FFTDirection::FORWARD) == SHAFFT_ERR_PLAN_NOT_INIT,
         "execute on a destroyed handle -> ERR_PLAN_NOT_INIT");
  report(shafft::fft1dPlan(*h, N, shafft::FFTType::C2C, MPI_COMM_NULL, nullptr, nullptr) ==
             SHAFFT_ERR_INVALID_COMM, "null communicator -> ERR_INVALID_COMM");
  delete h;
  report(hipStreamDestroy(s) == hipSuccess, "stream destroyed");
}

}  // namespace

int main(int argc, char** argv) {
  MPI_Init(&argc, &argv);
  int size = 1;
  MPI_Comm_size(MPI_COMM_WORLD, &size);
  if (argc > 1 && std::strcmp(argv[1], "multirank") == 0) {
    shafft::FFT1DHandle h;
    size_t a, b, c, d, e;
    const bool refused =
        shafft::fft1dPlan(h, 1024, shafft::FFTType::C2C, MPI_COMM_WORLD, nullptr, nullptr) ==
            SHAFFT_ERR_NOT_IMPL &&
        shafft::fft1dQueryLayout(1024, a, b, c, d, e, MPI_COMM_WORLD) == SHAFFT_ERR_NOT_IMPL &&
        h.internal == nullptr;
    report(size > 1 && refused, "FFT1D on a communicator of more than one rank -> ERR_NOT_IMPL");
    nd_partial_cases();
  } else {
    nd_partial_cases();
    streams();
    fft1d();
  }
  std::printf("%s: %d failure(s)\n", failures ? "FAILED" : "OK", failures);
  MPI_Finalize();
  return failures ? 1 : 0;
}
