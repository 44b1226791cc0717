// One-rank run of the C++ class API (include/shafft_amd.hpp) in the call
// sequences of the reference's C++ tests (tests/integration/cpp/
// test_roundtrip.cpp, test_FFT1D_roundtrip.cpp, test_correctness.cpp):
// FFTND / FFT1D init -> plan -> setBuffers -> forward -> normalize ->
// backward -> normalize, layout and axis queries, configuration queries,
// buffer helpers and moved-from plans.  Prints PASS/FAIL lines and exits with
// the number of failures; "host" runs only the checks without device calls.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "shafft_amd.hpp"

static int g_fail = 0;

static void check(const char* name, bool ok) {
  if (ok) {
    std::printf("PASS %s\n", name);
  } else {
    char msg[256] = {0};
    shafftLastErrorMessage(msg, (int)sizeof msg);
    std::printf("FAIL %s (%s)\n", name, msg);
    ++g_fail;
  }
  std::fflush(stdout);
}

static void host_checks() {
  shafft::FFTND nd;
  check("FFTND init", nd.init({1, 1, 1}, {32, 24, 10}, shafft::FFTType::C2C, shafft::kOneRank) == 0);
  std::vector<size_t> sub, off;
  std::vector<int> ca, da;
  check("FFTND sizes", nd.allocSize() == 32 * 24 * 10 && nd.globalSize() == 32 * 24 * 10 &&
                           nd.ndim() == 3 && nd.fftType() == shafft::FFTType::C2C && nd.isActive());
  check("FFTND getLayout", nd.getLayout(sub, off, shafft::TensorLayout::INITIAL) == 0 &&
                               sub == std::vector<size_t>{32, 24, 10} &&
                               off == std::vector<size_t>{0, 0, 0});
  check("FFTND getAxes", nd.getAxes(ca, da) == 0 && ca == std::vector<int>{0, 1, 2} && da.empty());
  shafft::FFTND moved(std::move(nd));
  check("moved-from FFTND holds no plan", nd.handle() == nullptr && moved.handle() != nullptr &&
                                              moved.allocSize() == 32 * 24 * 10);
  shafft::FFTND bad;
  check("FFTND init mismatched commDims -> DIM_MISMATCH",
        bad.init({1, 1}, {8, 8, 8}, shafft::FFTType::Z2Z, shafft::kOneRank) == SHAFFT_ERR_DIM_MISMATCH);
  check("FFTND init comm != 0 -> error",
        bad.init({1, 1}, {8, 8}, shafft::FFTType::Z2Z, 5) != 0);

  shafft::FFT1D od;
  size_t ln = 0, ls = 9;
  check("configuration1D", shafft::configuration1D(1000, ln, ls, shafft::FFTType::Z2Z, shafft::kOneRank) == 0 &&
                               ln == 1000 && ls == 0);
  check("FFT1D init", od.init(1000, ln, ls, shafft::FFTType::Z2Z, shafft::kOneRank) == 0 &&
                          od.allocSize() == 1000 && od.ndim() == 1);
  check("FFT1D getAxes: ca {} da {0}", od.getAxes(ca, da) == 0 && ca.empty() && da == std::vector<int>{0});

  std::vector<int> grid;
  int nda = -1, cs = 0;
  check("configurationND", shafft::configurationND({64, 64, 32}, shafft::FFTType::C2C, grid, nda, sub, off, cs,
                                                   shafft::DecompositionStrategy::MAXIMIZE_NDA, 0,
                                                   shafft::kOneRank) == 0 &&
                               nda == 0 && cs == 1 && sub == std::vector<size_t>{64, 64, 32});
  shafft_nd_config_t cfg;
  std::memset(&cfg, 0, sizeof cfg);
  const size_t shape[2] = {96, 210};
  shafft::FFTND fromcfg;
  check("FFTND init from a config object",
        shafftConfigNDInit(&cfg, 2, shape, SHAFFT_Z2Z, nullptr, 0, SHAFFT_MAXIMIZE_NDA,
                           SHAFFT_LAYOUT_REDISTRIBUTED, 0, shafft::kOneRank) == 0 &&
            fromcfg.init(cfg) == 0 && fromcfg.allocSize() == 96 * 210 && fromcfg.ndim() == 2 &&
            fromcfg.fftType() == shafft::FFTType::Z2Z);
  shafftConfigNDRelease(&cfg);
  shafft::Comm comm = 7;
  check("getCommunicator", fromcfg.getCommunicator(&comm) == 0 && comm == shafft::kOneRank);

  const shafft::Version v = shafft::getVersion();
  check("version and backend", v.major >= 1 && std::strchr(shafft::getVersionString(), '.') &&
                                   std::strlen(shafft::getBackendName()) > 0);
}

template <typename T>
static double nd_roundtrip(std::vector<size_t> dims, shafft::FFTType type) {
  using C = std::complex<T>;
  shafft::FFTND p0;
  if (p0.init(std::vector<int>(dims.size(), 1), dims, type, shafft::kOneRank) || p0.plan()) return 1e9;
  shafft::FFTND p = std::move(p0);  // a planned plan survives a move
  const size_t n = p.allocSize();
  C *data = nullptr, *work = nullptr;
  if (shafft::allocBuffer(n, &data) || shafft::allocBuffer(n, &work)) return 1e9;
  std::mt19937_64 rng(42);
  std::uniform_real_distribution<double> u(-1, 1);
  std::vector<C> x(n), y(n);
  for (auto& e : x) e = C((T)u(rng), (T)u(rng));
  int rc = shafft::copyToBuffer(data, x.data(), n) | p.setBuffers(data, work);
  rc |= p.execute(shafft::FFTDirection::FORWARD) | p.normalize();
  rc |= p.execute(shafft::FFTDirection::BACKWARD) | p.normalize();
  C *d = nullptr, *w = nullptr;
  rc |= p.getBuffers(&d, &w) | shafft::copyFromBuffer(y.data(), d, n);
  double num = 0, den = 0;
  for (size_t i = 0; i < n; ++i) {
    num += std::norm(std::complex<double>(y[i]) - std::complex<double>(x[i]));
    den += std::norm(std::complex<double>(x[i]));
  }
  rc |= shafft::freeBuffer(data) | shafft::freeBuffer(work);
  return rc ? 1e9 : std::sqrt(num / den);
}

template <typename T>
static double od_delta(size_t N, shafft::FFTType type) {
  using C = std::complex<T>;
  shafft::FFT1D p;
  size_t ln = 0, ls = 0;
  if (shafft::configuration1D(N, ln, ls, type, shafft::kOneRank) || p.init(N, ln, ls, type, shafft::kOneRank) ||
      p.plan())
    return 1e9;
  const size_t n = p.allocSize();
  C *data = nullptr, *work = nullptr;
  if (shafft::allocBuffer(n, &data) || shafft::allocBuffer(n, &work)) return 1e9;
  std::vector<C> x(n, C(0, 0)), y(n);
  x[0] = C(1, 0);
  int rc = shafft::copyToBuffer(data, x.data(), n) | p.setBuffers(data, work);
  rc |= p.execute(shafft::FFTDirection::FORWARD) | p.normalize();
  C *d = nullptr, *w = nullptr;
  rc |= p.getBuffers(&d, &w) | shafft::copyFromBuffer(y.data(), d, n);
  const double c = 1.0 / std::sqrt((double)N);
  double worst = 0;
  for (size_t i = 0; i < N; ++i) worst = std::max(worst, std::abs(std::complex<double>(y[i]) - c) / c);
  rc |= shafft::freeBuffer(data) | shafft::freeBuffer(work);
  return rc ? 1e9 : worst;
}

int main(int argc, char** argv) {
  host_checks();
  if (!(argc > 1 && std::strcmp(argv[1], "host") == 0)) {
    double e = nd_roundtrip<float>({32, 32, 32}, shafft::FFTType::C2C);
    std::printf("  32^3 complex64 round trip %.3e\n", e);
    check("FFTND round trip 32^3 complex64", e <= 1e-5);
    e = nd_roundtrip<double>({64, 64, 32}, shafft::FFTType::Z2Z);
    std::printf("  64x64x32 complex128 round trip (moved plan) %.3e\n", e);
    check("FFTND round trip 64x64x32 complex128, moved plan", e <= 1e-12);
    e = od_delta<float>(256, shafft::FFTType::C2C);
    check("FFT1D delta N=256 complex64", e <= 1e-5);
    e = od_delta<double>(1009, shafft::FFTType::Z2Z);
    check("FFT1D delta N=1009 complex128", e <= 1e-12);
  }
  std::printf("%d check(s) failed\n", g_fail);
  return g_fail;
}
