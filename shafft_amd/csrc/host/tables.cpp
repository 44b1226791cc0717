#include "tables.hpp"

#include <cmath>
#include <algorithm>
#include <stdexcept>
#include <system_error>
#include <thread>
#include <utility>

namespace sa {

static const long double kTwoPi = 6.283185307179586476925286766559005768L;

// cos and sin of 2 pi num / den for 0 <= num < den.  Exact integer
// comparisons fold the angle into [0, pi/4] before the long double call.
static void cs_turn(int64_t num, int64_t den, long double& c, long double& s) {
  if (2 * num > den) {  // 2pi - x
    cs_turn(den - num, den, c, s);
    s = -s;
    return;
  }
  if (4 * num > den) {  // pi - x
    cs_turn(den - 2 * num, 2 * den, c, s);
    c = -c;
    return;
  }
  if (8 * num > den) {  // pi/2 - x
    cs_turn(den - 4 * num, 4 * den, c, s);
    std::swap(c, s);
    return;
  }
  const long double a = kTwoPi * (long double)num / (long double)den;
  c = cosl(a);
  s = sinl(a);
}

cld root_fwd(int64_t m, int64_t L) {
  m %= L;
  if (m < 0) m += L;
  long double c, s;
  cs_turn(m, L, c, s);
  return cld(c, -s);
}

std::vector<cld> stage_twiddles(const std::vector<int>& radices, std::vector<int>& offsets) {
  std::vector<cld> t;
  offsets.clear();
  int64_t Ns = 1;
  for (int R : radices) {
    offsets.push_back((int)t.size());
    if (Ns > 1) {
      const int64_t L = Ns * R;
      for (int64_t rem = 0; rem < Ns; ++rem)
        for (int r = 1; r < R; ++r) t.push_back(root_fwd(rem * r, L));
    }
    Ns *= R;
  }
  if (t.empty()) t.push_back(cld(1, 0));
  return t;
}

void two_level(int64_t N, int& shift, std::vector<cld>& lo, std::vector<cld>& hi) {
  int bits = 0;
  while ((int64_t(1) << bits) < N) ++bits;
  shift = (bits + 1) / 2;
  if (shift < 1) shift = 1;
  const int64_t L = int64_t(1) << shift;
  lo.resize(L);
  for (int64_t j = 0; j < L; ++j) lo[j] = root_fwd(j, N);
  const int64_t H = (N + L - 1) / L;
  hi.resize(H);
  for (int64_t i = 0; i < H; ++i) hi[i] = root_fwd(i * L, N);
}

namespace {

// Complex product by the textbook formula (std::complex's operator* goes
// through the C library's inf/nan-aware routine, several times slower).
inline cld mul(const cld& a, const cld& b) {
  return cld(a.real() * b.real() - a.imag() * b.imag(), a.real() * b.imag() + a.imag() * b.real());
}

// f(begin, end) over [0, n) on up to 16 threads, one thread per `grain`
// items at least (the tables of long transforms take seconds on one core).
template <class F>
void parallel_for(int64_t n, int64_t grain, const F& f) {
  const unsigned hw = std::thread::hardware_concurrency();
  const int64_t nt = std::min<int64_t>(std::min<int64_t>(hw ? hw : 1, 16), n / grain);
  if (nt <= 1) {
    f(int64_t(0), n);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(nt);
  int64_t t = 0;
  try {
    for (; t < nt; ++t) th.emplace_back([&f, n, nt, t] { f(n * t / nt, n * (t + 1) / nt); });
  } catch (const std::system_error&) {
    f(n * t / nt, n);  // no more threads: the rest on this one
  }
  for (auto& x : th) x.join();
}

uint64_t bit_reverse(uint64_t x, int bits) {
  x = ((x >> 1) & 0x5555555555555555ull) | ((x & 0x5555555555555555ull) << 1);
  x = ((x >> 2) & 0x3333333333333333ull) | ((x & 0x3333333333333333ull) << 2);
  x = ((x >> 4) & 0x0F0F0F0F0F0F0F0Full) | ((x & 0x0F0F0F0F0F0F0F0Full) << 4);
  x = __builtin_bswap64(x);
  return bits ? x >> (64 - bits) : 0;
}

}  // namespace

void fft_ld(std::vector<cld>& a) {
  const int64_t M = (int64_t)a.size();
  if (M & (M - 1)) throw std::invalid_argument("fft_ld: size not a power of two");
  if (M < 2) return;
  int bits = 0;
  while ((int64_t(1) << bits) < M) ++bits;
  constexpr int64_t kGrain = int64_t(1) << 15;
  parallel_for(M, kGrain, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      const int64_t j = (int64_t)bit_reverse((uint64_t)i, bits);
      if (i < j) std::swap(a[i], a[j]);
    }
  });
  // W_M^k for k < M/2; level `len` uses W_len^k = W_M^{k M/len} (the same
  // value bit for bit: cs_turn reduces num/den exactly and M/len is a power
  // of two).
  std::vector<cld> w(M / 2);
  parallel_for(M / 2, kGrain, [&](int64_t b, int64_t e) {
    for (int64_t k = b; k < e; ++k) w[k] = root_fwd(k, M);
  });
  for (int l = 1; l <= bits; ++l) {
    const int64_t h = int64_t(1) << (l - 1);
    const int sh = bits - l;  // w index = k << sh
    parallel_for(M / 2, kGrain, [&](int64_t b, int64_t e) {
      for (int64_t t = b; t < e; ++t) {
        const int64_t k = t & (h - 1);
        const int64_t i = ((t >> (l - 1)) << l) + k;
        const cld u = a[i], v = mul(a[i + h], w[k << sh]);
        a[i] = u + v;
        a[i + h] = u - v;
      }
    });
  }
}

void bluestein_tables(int64_t n, int64_t M, int sign, std::vector<cld>& chirp,
                      std::vector<cld>& H) {
  constexpr int64_t kGrain = int64_t(1) << 15;
  chirp.resize(n);
  const int64_t twon = 2 * n;
  parallel_for(n, kGrain, [&](int64_t b, int64_t e) {
    for (int64_t k = b; k < e; ++k) {
      // exp(sign i pi q / n) = exp(sign 2 pi i q / (2n)); root_fwd gives exp(-2 pi i m/L)
      const int64_t q = (int64_t)(((__int128)k * k) % twon);
      const cld r = root_fwd(q, twon);
      chirp[k] = sign < 0 ? r : std::conj(r);
    }
  });
  std::vector<cld> h(M, cld(0, 0));
  h[0] = std::conj(chirp[0]);
  for (int64_t m = 1; m < n; ++m) {
    h[m] = std::conj(chirp[m]);
    h[M - m] = std::conj(chirp[m]);
  }
  fft_ld(h);
  const long double inv = 1.0L / (long double)M;
  H.resize(M);
  parallel_for(M, kGrain, [&](int64_t b, int64_t e) {
    for (int64_t k = b; k < e; ++k) H[k] = h[k] * inv;
  });
}

void bluestein_tables_pair(int64_t n, int64_t M, std::vector<cld> chirp[2], std::vector<cld> H[2]) {
  bluestein_tables(n, M, -1, chirp[0], H[0]);
  chirp[1].resize(n);
  for (int64_t k = 0; k < n; ++k) chirp[1][k] = std::conj(chirp[0][k]);
  H[1].resize(M);
  for (int64_t k = 0; k < M; ++k) H[1][k] = std::conj(H[0][k]);
}

}  // namespace sa
