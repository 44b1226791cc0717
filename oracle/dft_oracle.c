/* FP64 DFT oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A CPU restatement of what SHAFFT's FFTND/FFT1D compute at one rank: the
 * unnormalised multi-axis DFT over the transformed axes of a row-major
 * complex tensor, forward sign -1 (tests/unit/test_fft_partial.cpp:109-119,
 * tests/unit/test_fft_highdim.cpp:53-58), backward sign +1, and the
 * normalisation factor (1/sqrt(prod dims))^e (src/cpp/shafft_internal.cpp:
 * 129-188).  Only tests, smoke() and the benchmark's CPU-baseline leg use it,
 * and only to check results or as the timed CPU baseline -- never as a product
 * path.
 *
 * oracle_dft_axis: direct O(n^2)-per-line DFT ("dft_partial" in
 *   tests/unit/test_fft_partial.cpp:43-124) with the twiddle index
 *   (j*k) mod n reduced exactly in integer arithmetic, roots evaluated in
 *   long double, sums accumulated in long double.
 * oracle_fft_axis: the same transform by an iterative radix-2 FFT in
 *   double for power-of-two n (CPU baseline; pinned against the direct DFT
 *   by the tests), with OpenMP over lines.
 *
 * oracle_uniform / oracle_gaussian: the synthetic input recipe
 *   (std::mt19937_64 seeded per workload, uniform [-1,1) or Box-Muller
 *   Gaussian re/im per element).
 *
 * Complex data is interleaved (re, im) doubles.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

static const long double kTwoPiL = 6.283185307179586476925286766559005768L;

static void line_geometry(int ndim, const int64_t* dims, int axis, int64_t* outer, int64_t* n,
                          int64_t* inner) {
  int64_t o = 1, s = 1;
  for (int i = 0; i < axis; ++i) o *= dims[i];
  for (int i = axis + 1; i < ndim; ++i) s *= dims[i];
  *outer = o;
  *n = dims[axis];
  *inner = s;
}

int oracle_dft_axis(const double* in, double* out, int ndim, const int64_t* dims, int axis,
                    int sign) {
  if (axis < 0 || axis >= ndim) return 1;
  int64_t outer, n, inner;
  line_geometry(ndim, dims, axis, &outer, &n, &inner);
  long double* wr = (long double*)malloc(sizeof(long double) * n);
  long double* wi = (long double*)malloc(sizeof(long double) * n);
  if (!wr || !wi) return 2;
  for (int64_t m = 0; m < n; ++m) {
    const long double a = kTwoPiL * (long double)m / (long double)n;
    wr[m] = cosl(a);
    wi[m] = (sign < 0 ? -1.0L : 1.0L) * sinl(a);
  }
#pragma omp parallel for collapse(2) schedule(static)
  for (int64_t o = 0; o < outer; ++o)
    for (int64_t s = 0; s < inner; ++s) {
      const int64_t base = o * n * inner + s;
      for (int64_t k = 0; k < n; ++k) {
        long double ar = 0, ai = 0;
        for (int64_t j = 0; j < n; ++j) {
          const int64_t m = (int64_t)(((__int128)j * k) % n);
          const long double xr = in[2 * (base + j * inner)];
          const long double xi = in[2 * (base + j * inner) + 1];
          ar += xr * wr[m] - xi * wi[m];
          ai += xr * wi[m] + xi * wr[m];
        }
        out[2 * (base + k * inner)] = (double)ar;
        out[2 * (base + k * inner) + 1] = (double)ai;
      }
    }
  free(wr);
  free(wi);
  return 0;
}

/* Iterative radix-2 DIT FFT of one strided line (n a power of two). */
static void fft_line_pow2(double* x, int64_t n, int64_t stride, int sign, const double* wr,
                          const double* wi, double* tmp) {
  for (int64_t i = 0; i < n; ++i) {
    tmp[2 * i] = x[2 * i * stride];
    tmp[2 * i + 1] = x[2 * i * stride + 1];
  }
  for (int64_t i = 1, j = 0; i < n; ++i) {
    int64_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      double t0 = tmp[2 * i], t1 = tmp[2 * i + 1];
      tmp[2 * i] = tmp[2 * j];
      tmp[2 * i + 1] = tmp[2 * j + 1];
      tmp[2 * j] = t0;
      tmp[2 * j + 1] = t1;
    }
  }
  for (int64_t len = 2; len <= n; len <<= 1) {
    const int64_t h = len >> 1, step = n / len;
    for (int64_t i = 0; i < n; i += len)
      for (int64_t k = 0; k < h; ++k) {
        const double c = wr[k * step], s = sign < 0 ? -wi[k * step] : wi[k * step];
        double* u = tmp + 2 * (i + k);
        double* v = tmp + 2 * (i + k + h);
        const double vr = v[0] * c - v[1] * s, vi = v[0] * s + v[1] * c;
        v[0] = u[0] - vr;
        v[1] = u[1] - vi;
        u[0] += vr;
        u[1] += vi;
      }
  }
  for (int64_t i = 0; i < n; ++i) {
    x[2 * i * stride] = tmp[2 * i];
    x[2 * i * stride + 1] = tmp[2 * i + 1];
  }
}

/* In-place FFT along `axis` for a power-of-two extent; returns 3 otherwise. */
int oracle_fft_axis(double* x, int ndim, const int64_t* dims, int axis, int sign) {
  int64_t outer, n, inner;
  line_geometry(ndim, dims, axis, &outer, &n, &inner);
  if (n & (n - 1)) return 3;
  double* wr = (double*)malloc(sizeof(double) * n);
  double* wi = (double*)malloc(sizeof(double) * n);
  for (int64_t m = 0; m < n; ++m) {
    const long double a = kTwoPiL * (long double)m / (long double)n;
    wr[m] = (double)cosl(a);
    wi[m] = (double)sinl(a);
  }
#pragma omp parallel
  {
    double* tmp = (double*)malloc(sizeof(double) * 2 * n);
#pragma omp for collapse(2) schedule(static)
    for (int64_t o = 0; o < outer; ++o)
      for (int64_t s = 0; s < inner; ++s)
        fft_line_pow2(x + 2 * (o * n * inner + s), n, inner, sign, wr, wi, tmp);
    free(tmp);
  }
  free(wr);
  free(wi);
  return 0;
}

/* x *= f over count complex elements (the reference's normalize loop,
 * src/cpp/fft/fftw_method/common/normalize.cpp:6-32). */
void oracle_scale(double* x, int64_t count, double f) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < 2 * count; ++i) x[i] *= f;
}

int oracle_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* std::mt19937_64 (Matsumoto & Nishimura 2000, 64-bit parameters) and the
 * input recipe: per element re then im, u = (x >> 11) * 2^-53, value 2u-1. */
typedef struct {
  uint64_t mt[312];
  int idx;
} mt64_t;

static void mt64_seed(mt64_t* s, uint64_t seed) {
  s->mt[0] = seed;
  for (int i = 1; i < 312; ++i)
    s->mt[i] = 6364136223846793005ULL * (s->mt[i - 1] ^ (s->mt[i - 1] >> 62)) + (uint64_t)i;
  s->idx = 312;
}

static uint64_t mt64_next(mt64_t* s) {
  if (s->idx >= 312) {
    for (int i = 0; i < 312; ++i) {
      const uint64_t x = (s->mt[i] & 0xFFFFFFFF80000000ULL) | (s->mt[(i + 1) % 312] & 0x7FFFFFFFULL);
      uint64_t xa = x >> 1;
      if (x & 1ULL) xa ^= 0xB5026F5AA96619E9ULL;
      s->mt[i] = s->mt[(i + 156) % 312] ^ xa;
    }
    s->idx = 0;
  }
  uint64_t y = s->mt[s->idx++];
  y ^= (y >> 29) & 0x5555555555555555ULL;
  y ^= (y << 17) & 0x71D67FFFEDA60000ULL;
  y ^= (y << 37) & 0xFFF7EEE000000000ULL;
  y ^= y >> 43;
  return y;
}

void oracle_uniform(double* out, int64_t count, uint64_t seed) {
  mt64_t s;
  mt64_seed(&s, seed);
  for (int64_t i = 0; i < 2 * count; ++i) {
    const double u = (double)(mt64_next(&s) >> 11) * (1.0 / 9007199254740992.0);
    out[i] = 2.0 * u - 1.0;
  }
}

/* Gaussian input recipe (Box-Muller on consecutive draws): per element
 * u1 = ((x0 >> 11) + 1) * 2^-53 in (0, 1], u2 = (x1 >> 11) * 2^-53 in
 * [0, 1), r = sqrt(-2 ln u1), re = r cos(2 pi u2), im = r sin(2 pi u2). */
void oracle_gaussian(double* out, int64_t count, uint64_t seed) {
  mt64_t s;
  mt64_seed(&s, seed);
  const double two_pi = 6.283185307179586476925286766559;
  for (int64_t i = 0; i < count; ++i) {
    const double u1 = (double)((mt64_next(&s) >> 11) + 1) * (1.0 / 9007199254740992.0);
    const double u2 = (double)(mt64_next(&s) >> 11) * (1.0 / 9007199254740992.0);
    const double r = sqrt(-2.0 * log(u1));
    out[2 * i] = r * cos(two_pi * u2);
    out[2 * i + 1] = r * sin(two_pi * u2);
  }
}

/* First raw 64-bit outputs (for pinning the generator against the
 * published std::mt19937_64 check value). */
void oracle_mt64_raw(uint64_t* out, int64_t count, uint64_t seed) {
  mt64_t s;
  mt64_seed(&s, seed);
  for (int64_t i = 0; i < count; ++i) out[i] = mt64_next(&s);
}
