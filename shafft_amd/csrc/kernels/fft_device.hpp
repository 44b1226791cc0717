// Device-side building blocks for the gfx950 FFT kernels: complex arithmetic,
// fixed-radix DFT butterflies (2..10,12,14,15,16,18,20,21,24,25,28), and 32-bit division by a
// run-time invariant.
//
// Conventions
//   * SIGN = -1 is the forward transform  X[k] = sum_n x[n] exp(-2 pi i n k / N)
//     (the sign SHAFFT uses: tests/unit/test_fft_partial.cpp:109-119),
//     SIGN = +1 is the unnormalised backward transform.
//   * Twiddle tables always hold forward roots exp(-2 pi i m / L); the
//     backward kernels multiply by their conjugate.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sa {

template <typename R>
struct alignas(2 * sizeof(R)) cpx {
  R x, y;
};

template <typename R>
__device__ __forceinline__ cpx<R> cadd(cpx<R> a, cpx<R> b) {
  return {a.x + b.x, a.y + b.y};
}
template <typename R>
__device__ __forceinline__ cpx<R> csub(cpx<R> a, cpx<R> b) {
  return {a.x - b.x, a.y - b.y};
}
template <typename R>
__device__ __forceinline__ cpx<R> cmul(cpx<R> a, cpx<R> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename R>
__device__ __forceinline__ cpx<R> cscale(cpx<R> a, R s) {
  return {a.x * s, a.y * s};
}
template <typename R>
__device__ __forceinline__ cpx<R> cconj(cpx<R> a) {
  return {a.x, -a.y};
}
// a * w for the forward direction, a * conj(w) for the backward one.
template <int SIGN, typename R>
__device__ __forceinline__ cpx<R> cmul_dir(cpx<R> a, cpx<R> w) {
  if constexpr (SIGN < 0) {
    return {a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x};
  } else {
    return {a.x * w.x + a.y * w.y, a.y * w.x - a.x * w.y};
  }
}
// a * exp(SIGN * i * pi / 2)
template <int SIGN, typename R>
__device__ __forceinline__ cpx<R> rot90(cpx<R> a) {
  if constexpr (SIGN < 0) {
    return {a.y, -a.x};
  } else {
    return {-a.y, a.x};
  }
}
// a * (c + SIGN * i * s) with compile-time-known constants c, s.
template <int SIGN, typename R>
__device__ __forceinline__ cpx<R> crot(cpx<R> a, R c, R s) {
  const R ss = SIGN < 0 ? -s : s;
  return {a.x * c - a.y * ss, a.x * ss + a.y * c};
}

#ifdef __HIPCC__
// Copies n complex values from global memory into LDS with the whole
// workgroup.  Each thread issues U loads before its first LDS store (a plain
// strided loop waits for every load in turn: one L2 round trip per element
// per thread); out-of-range slots load element n-1 and are not stored, so no
// load is predicated.
template <typename R, int U = 8>
__device__ __forceinline__ void copy_to_lds(cpx<R>* __restrict__ dst, const cpx<R>* __restrict__ src,
                                            int n) {
  const int nt = (int)blockDim.x;
  for (int base = (int)threadIdx.x; base < n; base += U * nt) {
    cpx<R> v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = src[min(base + u * nt, n - 1)];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (base + u * nt < n) dst[base + u * nt] = v[u];
  }
}
#endif

// ---------------------------------------------------------------------------
// Butterflies: in-register DFT of length RAD, X[q] = sum_r v[r] w_RAD^{SIGN r q}.
// ---------------------------------------------------------------------------

template <int SIGN, typename R>
__device__ __forceinline__ void dft2(cpx<R>* v) {
  cpx<R> a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <int SIGN, typename R>
__device__ __forceinline__ void dft4(cpx<R>& a0, cpx<R>& a1, cpx<R>& a2, cpx<R>& a3) {
  cpx<R> t0 = cadd(a0, a2), t1 = csub(a0, a2);
  cpx<R> t2 = cadd(a1, a3), t3 = rot90<SIGN>(csub(a1, a3));
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  a1 = cadd(t1, t3);
  a3 = csub(t1, t3);
}

template <int SIGN, typename R>
__device__ __forceinline__ void dft4(cpx<R>* v) {
  dft4<SIGN>(v[0], v[1], v[2], v[3]);
}

template <int SIGN, typename R>
__device__ __forceinline__ void dft8(cpx<R>* v) {
  constexpr R c = R(0.7071067811865475244008444);
  // even / odd radix-4 halves
  cpx<R> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  cpx<R> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4<SIGN>(e0, e1, e2, e3);
  dft4<SIGN>(o0, o1, o2, o3);
  o1 = crot<SIGN>(o1, c, c);
  o2 = rot90<SIGN>(o2);
  o3 = crot<SIGN>(o3, -c, c);
  v[0] = cadd(e0, o0);
  v[4] = csub(e0, o0);
  v[1] = cadd(e1, o1);
  v[5] = csub(e1, o1);
  v[2] = cadd(e2, o2);
  v[6] = csub(e2, o2);
  v[3] = cadd(e3, o3);
  v[7] = csub(e3, o3);
}

template <int SIGN, typename R>
__device__ __forceinline__ void dft16(cpx<R>* v) {
  constexpr R c1 = R(0.9238795325112867561281832);  // cos(pi/8)
  constexpr R s1 = R(0.3826834323650897717284600);  // sin(pi/8)
  constexpr R c2 = R(0.7071067811865475244008444);  // cos(pi/4)
  // Column DFT4s over r' for each r0 (r = 4 r' + r0).
#pragma unroll
  for (int r0 = 0; r0 < 4; ++r0) dft4<SIGN>(v[r0], v[r0 + 4], v[r0 + 8], v[r0 + 12]);
  // v[r0 + 4 q0] now holds F_{r0}[q0]; twiddle by w16^{r0 q0}.
  v[1 + 4 * 1] = crot<SIGN>(v[1 + 4 * 1], c1, s1);    // m=1
  v[1 + 4 * 2] = crot<SIGN>(v[1 + 4 * 2], c2, c2);    // m=2
  v[1 + 4 * 3] = crot<SIGN>(v[1 + 4 * 3], s1, c1);    // m=3
  v[2 + 4 * 1] = crot<SIGN>(v[2 + 4 * 1], c2, c2);    // m=2
  v[2 + 4 * 2] = rot90<SIGN>(v[2 + 4 * 2]);           // m=4
  v[2 + 4 * 3] = crot<SIGN>(v[2 + 4 * 3], -c2, c2);   // m=6
  v[3 + 4 * 1] = crot<SIGN>(v[3 + 4 * 1], s1, c1);    // m=3
  v[3 + 4 * 2] = crot<SIGN>(v[3 + 4 * 2], -c2, c2);   // m=6
  v[3 + 4 * 3] = crot<SIGN>(v[3 + 4 * 3], -c1, -s1);  // m=9
  // Row DFT4s over r0 for each q0: X[q0 + 4 q1].
  cpx<R> out[16];
#pragma unroll
  for (int q0 = 0; q0 < 4; ++q0) {
    cpx<R> a0 = v[0 + 4 * q0], a1 = v[1 + 4 * q0], a2 = v[2 + 4 * q0], a3 = v[3 + 4 * q0];
    dft4<SIGN>(a0, a1, a2, a3);
    out[q0 + 0] = a0;
    out[q0 + 4] = a1;
    out[q0 + 8] = a2;
    out[q0 + 12] = a3;
  }
#pragma unroll
  for (int i = 0; i < 16; ++i) v[i] = out[i];
}

// Odd prime radix by the symmetric-pair formulation:
//   X[q]   = A_q + SIGN i B_q,  X[R-q] = A_q - SIGN i B_q
//   A_q = v0 + sum_r (v_r + v_{R-r}) cos(2 pi r q / R)
//   B_q =      sum_r (v_r - v_{R-r}) sin(2 pi r q / R)
template <int RAD>
struct OddConst;
template <>
struct OddConst<3> {
  static constexpr double c[2] = {1.0, -0.5};
  static constexpr double s[2] = {0.0, 0.8660254037844386467637232};
};
template <>
struct OddConst<5> {
  static constexpr double c[3] = {1.0, 0.3090169943749474241022934, -0.8090169943749474241022934};
  static constexpr double s[3] = {0.0, 0.9510565162951535721164393, 0.5877852522924731291687060};
};
template <>
struct OddConst<7> {
  static constexpr double c[4] = {1.0, 0.6234898018587335305250049, -0.2225209339563144042889026,
                                  -0.9009688679024191262361023};
  static constexpr double s[4] = {0.0, 0.7818314824680298087084445, 0.9749279121818236070181317,
                                  0.4338837391175581204757683};
};

template <int RAD, int SIGN, typename R>
__device__ __forceinline__ void dft_odd(cpx<R>* v) {
  constexpr int H = (RAD - 1) / 2;
  cpx<R> S[H + 1], D[H + 1];
  cpx<R> x0 = v[0];
  cpx<R> sum = x0;
#pragma unroll
  for (int r = 1; r <= H; ++r) {
    S[r] = cadd(v[r], v[RAD - r]);
    D[r] = csub(v[r], v[RAD - r]);
    sum = cadd(sum, S[r]);
  }
  v[0] = sum;
#pragma unroll
  for (int q = 1; q <= H; ++q) {
    cpx<R> A = x0, B = {R(0), R(0)};
#pragma unroll
    for (int r = 1; r <= H; ++r) {
      const int m = (r * q) % RAD;
      // cos(2 pi m/R) = c[min(m, R-m)], sin(2 pi m/R) = +-s[min(m, R-m)]
      const int mm = m <= H ? m : RAD - m;
      const R cm = R(OddConst<RAD>::c[mm]);
      const R sm = m <= H ? R(OddConst<RAD>::s[mm]) : -R(OddConst<RAD>::s[mm]);
      A.x += S[r].x * cm;
      A.y += S[r].y * cm;
      B.x += D[r].x * sm;
      B.y += D[r].y * sm;
    }
    // SIGN i B
    cpx<R> iB = SIGN < 0 ? cpx<R>{B.y, -B.x} : cpx<R>{-B.y, B.x};
    v[q] = cadd(A, iB);
    v[RAD - q] = csub(A, iB);
  }
}

template <int RAD, int SIGN, typename R>
__device__ __forceinline__ void dft(cpx<R>* v);

// Composite radix R = R1 * R2 with gcd(R1, R2) = 1 by the prime-factor
// (Good-Thomas) map: input n = (R2 n1 + R1 n2) mod R, output k with
// k = k1 (mod R1), k = k2 (mod R2); then
//   X[k] = sum_{n1,n2} x[n] w_R1^{n1 k1} w_R2^{n2 k2}
// (w_R^{nk} = w_R1^{n1 k} w_R2^{n2 k}), so the R2 inner length-R1 DFTs and R1
// outer length-R2 DFTs need no twiddles; both maps are compile-time.
template <int R1, int R2>
struct Crt {
  // the k in [0, R1 R2) with k % R1 == k1 and k % R2 == k2
  static constexpr int at(int k1, int k2) {
    for (int k = 0; k < R1 * R2; ++k)
      if (k % R1 == k1 && k % R2 == k2) return k;
    return -1;
  }
};

template <int R1, int R2, int SIGN, typename R>
__device__ __forceinline__ void dft_pfa(cpx<R>* v) {
  constexpr int RR = R1 * R2;
  cpx<R> y[R2][R1];
#pragma unroll
  for (int n2 = 0; n2 < R2; ++n2) {
    cpx<R> t[R1];
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) t[n1] = v[(R2 * n1 + R1 * n2) % RR];
    dft<R1, SIGN>(t);
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) y[n2][k1] = t[k1];
  }
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) {
    cpx<R> t[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) t[n2] = y[n2][k1];
    dft<R2, SIGN>(t);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) v[Crt<R1, R2>::at(k1, k2)] = t[k2];
  }
}

// Radix 9 = 3 x 3 by Cooley-Tukey: n = 3 n1 + n2, k = k1 + 3 k2,
//   X[k] = sum_{n2} w_9^{n2 k1} w_3^{n2 k2} sum_{n1} x[3 n1 + n2] w_3^{n1 k1},
// with the four non-trivial twiddles w_9^{m}, m = n2 k1 in {1, 2, 4}.
template <int SIGN, typename R>
__device__ __forceinline__ void dft9(cpx<R>* v) {
  constexpr R c1 = R(0.7660444431189780352023927), s1 = R(0.6427876096865393263226434);
  constexpr R c2 = R(0.1736481776669303488517166), s2 = R(0.9848077530122080593667430);
  constexpr R c4 = R(-0.9396926207859083840541093), s4 = R(0.3420201433256687330440996);
  cpx<R> y[3][3];
#pragma unroll
  for (int n2 = 0; n2 < 3; ++n2) {
    cpx<R> t[3] = {v[n2], v[3 + n2], v[6 + n2]};
    dft<3, SIGN>(t);
    y[n2][0] = t[0];
    y[n2][1] = t[1];
    y[n2][2] = t[2];
  }
  y[1][1] = crot<SIGN>(y[1][1], c1, s1);  // m = 1
  y[1][2] = crot<SIGN>(y[1][2], c2, s2);  // m = 2
  y[2][1] = crot<SIGN>(y[2][1], c2, s2);  // m = 2
  y[2][2] = crot<SIGN>(y[2][2], c4, s4);  // m = 4
#pragma unroll
  for (int k1 = 0; k1 < 3; ++k1) {
    cpx<R> t[3] = {y[0][k1], y[1][k1], y[2][k1]};
    dft<3, SIGN>(t);
    v[k1] = t[0];
    v[k1 + 3] = t[1];
    v[k1 + 6] = t[2];
  }
}

// Radix 25 = 5 x 5 by Cooley-Tukey (n = 5 n1 + n2, k = k1 + 5 k2) with the
// sixteen twiddles w_25^m, m = n2 k1 in 1..16: cos / sin(2 pi m / 25).
struct Tw25 {
  static constexpr double c[17] = {
      1.0, 0.9685831611286311194902, 0.8763066800438635873081,
      0.7289686274214115231467, 0.5358267949789966182713, 0.3090169943749474241023,
      0.06279051952931337607618, -0.1873813145857246305426, -0.4257792915650726488625,
      -0.6374239897486897101767, -0.8090169943749474241023, -0.9297764858882514036609,
      -0.9921147013144778310498, -0.9921147013144778310498, -0.9297764858882514036609,
      -0.8090169943749474241023, -0.6374239897486897101767};
  static constexpr double s[17] = {
      0.0, 0.2486898871648547882423, 0.4817536741017152749872,
      0.6845471059286886737323, 0.8443279255020150785486, 0.9510565162951535721164,
      0.9980267284282715619523, 0.9822872507286886810856, 0.9048270524660195277137,
      0.770513242775789230803, 0.5877852522924731291687, 0.3681245526846779591569,
      0.1253332335643042453731, -0.1253332335643042453731, -0.3681245526846779591569,
      -0.5877852522924731291687, -0.770513242775789230803};
};

template <int SIGN, typename R>
__device__ __forceinline__ void dft25(cpx<R>* v) {
  cpx<R> y[5][5];
#pragma unroll
  for (int n2 = 0; n2 < 5; ++n2) {
    cpx<R> t[5];
#pragma unroll
    for (int n1 = 0; n1 < 5; ++n1) t[n1] = v[5 * n1 + n2];
    dft<5, SIGN>(t);
#pragma unroll
    for (int k1 = 0; k1 < 5; ++k1) {
      const int m = n2 * k1;
      y[n2][k1] = m == 0 ? t[k1] : crot<SIGN>(t[k1], R(Tw25::c[m]), R(Tw25::s[m]));
    }
  }
#pragma unroll
  for (int k1 = 0; k1 < 5; ++k1) {
    cpx<R> t[5];
#pragma unroll
    for (int n2 = 0; n2 < 5; ++n2) t[n2] = y[n2][k1];
    dft<5, SIGN>(t);
#pragma unroll
    for (int k2 = 0; k2 < 5; ++k2) v[k1 + 5 * k2] = t[k2];
  }
}

template <int RAD, int SIGN, typename R>
__device__ __forceinline__ void dft(cpx<R>* v) {
  if constexpr (RAD == 2) dft2<SIGN>(v);
  else if constexpr (RAD == 4) dft4<SIGN>(v);
  else if constexpr (RAD == 8) dft8<SIGN>(v);
  else if constexpr (RAD == 16) dft16<SIGN>(v);
  else if constexpr (RAD == 6) dft_pfa<2, 3, SIGN>(v);
  else if constexpr (RAD == 10) dft_pfa<2, 5, SIGN>(v);
  else if constexpr (RAD == 12) dft_pfa<4, 3, SIGN>(v);
  else if constexpr (RAD == 9) dft9<SIGN>(v);
  else if constexpr (RAD == 14) dft_pfa<2, 7, SIGN>(v);
  else if constexpr (RAD == 15) dft_pfa<3, 5, SIGN>(v);
  else if constexpr (RAD == 18) dft_pfa<2, 9, SIGN>(v);
  else if constexpr (RAD == 20) dft_pfa<4, 5, SIGN>(v);
  else if constexpr (RAD == 21) dft_pfa<3, 7, SIGN>(v);
  else if constexpr (RAD == 24) dft_pfa<8, 3, SIGN>(v);
  else if constexpr (RAD == 25) dft25<SIGN>(v);
  else if constexpr (RAD == 28) dft_pfa<4, 7, SIGN>(v);
  else dft_odd<RAD, SIGN>(v);
}

// ---------------------------------------------------------------------------
// Division of a 31-bit unsigned value by a run-time invariant divisor using a
// multiply-high and a shift (Granlund & Montgomery, PLDI 1994).  Host side
// fills {d, m, s}; d == 1 is special-cased.
// ---------------------------------------------------------------------------
struct FastDiv {
  uint32_t d;
  uint32_t m;
  uint32_t s;
};

__device__ __forceinline__ uint32_t fdiv(uint32_t n, FastDiv f) {
  return f.d == 1 ? n : (__umulhi(n, f.m) >> f.s);
}

}  // namespace sa
