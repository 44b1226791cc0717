#include "tables.hpp"

#include <cmath>
#include <stdexcept>
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

void fft_ld(std::vector<cld>& a) {
  const int64_t M = (int64_t)a.size();
  if (M & (M - 1)) throw std::invalid_argument("fft_ld: size not a power of two");
  for (int64_t i = 1, j = 0; i < M; ++i) {
    int64_t bit = M >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  std::vector<cld> w;
  for (int64_t len = 2; len <= M; len <<= 1) {
    const int64_t h = len / 2;
    w.resize(h);
    for (int64_t k = 0; k < h; ++k) w[k] = root_fwd(k, len);
    for (int64_t i = 0; i < M; i += len)
      for (int64_t k = 0; k < h; ++k) {
        cld u = a[i + k], v = a[i + k + h] * w[k];
        a[i + k] = u + v;
        a[i + k + h] = u - v;
      }
  }
}

void bluestein_tables(int64_t n, int64_t M, int sign, std::vector<cld>& chirp,
                      std::vector<cld>& H) {
  chirp.resize(n);
  const int64_t twon = 2 * n;
  for (int64_t k = 0; k < n; ++k) {
    // exp(sign i pi q / n) = exp(sign 2 pi i q / (2n)); root_fwd gives exp(-2 pi i m/L)
    const int64_t q = (int64_t)(((__int128)k * k) % twon);
    const cld r = root_fwd(q, twon);
    chirp[k] = sign < 0 ? r : std::conj(r);
  }
  std::vector<cld> h(M, cld(0, 0));
  h[0] = std::conj(chirp[0]);
  for (int64_t m = 1; m < n; ++m) {
    h[m] = std::conj(chirp[m]);
    h[M - m] = std::conj(chirp[m]);
  }
  fft_ld(h);
  const long double inv = 1.0L / (long double)M;
  H.resize(M);
  for (int64_t k = 0; k < M; ++k) H[k] = h[k] * inv;
}

}  // namespace sa
