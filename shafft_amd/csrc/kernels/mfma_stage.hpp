// Radix-8 / radix-16 Stockham stage of a complex-float pass on the gfx950
// matrix cores, in split precision (opt-in).
//
// A radix-R butterfly is the dense product y = F_R x with the R x R DFT
// matrix.  In real arithmetic, with k = 2m + c indexing (Re, Im) of input m
// and i = 2p + c' indexing (Re, Im) of output p, it is a 2R x 2R matrix
//     A[2p][2m] = Wr   A[2p][2m+1] = -Wi   A[2p+1][2m] = Wi   A[2p+1][2m+1] = Wr
// with W = exp(SIGN 2 pi i p m / R), applied to 16 butterflies at once
// (the 16 columns of B).  Radix 16 uses v_mfma_f32_16x16x32 (K = 32, two
// 16-row output blocks), radix 8 v_mfma_f32_16x16x16 (K = 16, one block).
//
// Split precision: every operand is stored as hi + lo in 16-bit floats
// (hi = round(x), lo = round(x - hi)), and the product is accumulated in
// FP32 as A_hi B_hi + (A_hi B_lo + A_lo B_hi), dropping only A_lo B_lo.
//   MX = 1: bfloat16 halves (8-bit significands; range of FP32)
//   MX = 2: float16 halves (11-bit significands).  The lo halves are scaled
//           by 2^11 and each butterfly column by a power of two that brings
//           its largest input into [0.5, 1), so neither half leaves the
//           float16 range; the scales are undone in FP32.
// Lane layouts (checked on gfx950): A[i][k] in lane i + 16 (k / KPL),
// element k % KPL; B[k][j] in lane j + 16 (k / KPL); D[i][j] in lane
// j + 16 (i / 4), element i % 4.
#pragma once

#include "fft_device.hpp"

// Uses lp() and fdiv() from the including pass_tile.hpp / fft_device.hpp.

namespace sa {

typedef __bf16 mx_bf16x8 __attribute__((ext_vector_type(8)));
typedef short mx_s16x4 __attribute__((ext_vector_type(4)));
typedef _Float16 mx_h16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 mx_h16x4 __attribute__((ext_vector_type(4)));
typedef float mx_f32x4 __attribute__((ext_vector_type(4)));

// Operand vector of KPL 16-bit values for mode MX.
template <int MX, int KPL>
struct MxVec;
template <>
struct MxVec<1, 8> {
  typedef mx_bf16x8 type;
  static __device__ __forceinline__ void set(type& v, int e, float x) { v[e] = (__bf16)x; }
  static __device__ __forceinline__ float get(const type& v, int e) { return (float)v[e]; }
};
template <>
struct MxVec<1, 4> {
  typedef mx_s16x4 type;
  static __device__ __forceinline__ void set(type& v, int e, float x) {
    v[e] = __builtin_bit_cast(short, (__bf16)x);
  }
  // bfloat16 is the top half of a float: widen by shifting the bits.
  static __device__ __forceinline__ float get(const type& v, int e) {
    return __builtin_bit_cast(float, (uint32_t)(uint16_t)v[e] << 16);
  }
};
template <>
struct MxVec<2, 8> {
  typedef mx_h16x8 type;
  static __device__ __forceinline__ void set(type& v, int e, float x) { v[e] = (_Float16)x; }
  static __device__ __forceinline__ float get(const type& v, int e) { return (float)v[e]; }
};
template <>
struct MxVec<2, 4> {
  typedef mx_h16x4 type;
  static __device__ __forceinline__ void set(type& v, int e, float x) { v[e] = (_Float16)x; }
  static __device__ __forceinline__ float get(const type& v, int e) { return (float)v[e]; }
};

template <int MX, int KPL>
__device__ __forceinline__ mx_f32x4 mx_mma(const typename MxVec<MX, KPL>::type& a,
                                           const typename MxVec<MX, KPL>::type& b, mx_f32x4 c) {
  if constexpr (MX == 1 && KPL == 8) return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  if constexpr (MX == 2 && KPL == 8) return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
  if constexpr (MX == 1 && KPL == 4) return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
  if constexpr (MX == 2 && KPL == 4) return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
}

// cos(2 pi j / 16) for any integer j (correctly rounded doubles), by
// symmetry from cos(2 pi b / 16), b = 0..4; selects, no indexed memory.
__device__ __forceinline__ double mx_cos16(int j) {
  constexpr double c1 = 0.92387953251128675613, c2 = 0.70710678118654752440,
                   c3 = 0.38268343236508977173;
  const int q = j & 15;
  const int m = q <= 8 ? q : 16 - q;  // cos(2 pi (16 - q) / 16) = cos(2 pi q / 16)
  const bool neg = m > 4;             // cos(pi - x) = -cos(x)
  const int b = neg ? 8 - m : m;
  const double v = b == 0 ? 1.0 : b == 1 ? c1 : b == 2 ? c2 : b == 3 ? c3 : 0.0;
  return neg ? -v : v;
}

// The constant DFT operands of every (mode, radix, direction), built once
// by mx_table_init and loaded by the stages (each stage call built its own
// before: ~16 double-precision cosines, selects and conversions per lane per
// operand vector).  Slot (mode MX, radix index ri (0: 8, 1: 16), direction
// index si (0: forward), 16-row block b, half hl (0: hi, 1: lo), lane) holds
// the lane's KPL 16-bit values in its first 2 * KPL bytes.  Per translation
// unit (no relocatable device code); the batch launch runs mx_table_init on
// its stream before a split-precision batch kernel.
constexpr int kMxSlots = 2 * 2 * 2 * 2 * 2 * 64;
static __device__ __align__(16) uint4 g_mx_tab[kMxSlots];

__device__ __forceinline__ int mx_slot(int mx, int ri, int si, int b, int hl, int lane) {
  return (((((mx - 1) * 2 + ri) * 2 + si) * 2 + b) * 2 + hl) * 64 + lane;
}

// Operand vectors of one (MX, RAD, SIGN) for this lane.
template <int MX, int RAD, int SIGN>
__device__ __forceinline__ void mx_build(int lane, typename MxVec<MX, RAD == 16 ? 8 : 4>::type (&ahi)[RAD / 8],
                                         typename MxVec<MX, RAD == 16 ? 8 : 4>::type (&alo)[RAD / 8]) {
  constexpr int KPL = RAD == 16 ? 8 : 4;
  constexpr int MB = RAD / 8;
  constexpr float LO = MX == 2 ? 2048.f : 1.f;
  typedef MxVec<MX, KPL> V;
  const int j = lane & 15, h = lane >> 4;
#pragma unroll
  for (int b = 0; b < MB; ++b) {
    const int row = 16 * b + j, p = row >> 1, cp = row & 1;
#pragma unroll
    for (int e = 0; e < KPL; ++e) {
      const int k = KPL * h + e, m = k >> 1, cc = k & 1;
      const int q = (p * m) % RAD;
      const double wr = mx_cos16(q * (16 / RAD));
      const double wi = SIGN * mx_cos16(q * (16 / RAD) - 4);  // sin
      const double val = cp == 0 ? (cc == 0 ? wr : -wi) : (cc == 0 ? wi : wr);
      V::set(ahi[b], e, (float)val);
      V::set(alo[b], e, (float)((val - (double)V::get(ahi[b], e)) * LO));
    }
  }
}

template <int MX, int RAD, int SIGN>
__device__ __forceinline__ void mx_store(int lane) {
  typedef typename MxVec<MX, RAD == 16 ? 8 : 4>::type vec;
  vec ahi[RAD / 8], alo[RAD / 8];
  mx_build<MX, RAD, SIGN>(lane, ahi, alo);
  const int ri = RAD == 16 ? 1 : 0, si = SIGN < 0 ? 0 : 1;
#pragma unroll
  for (int b = 0; b < RAD / 8; ++b) {
    *reinterpret_cast<vec*>(&g_mx_tab[mx_slot(MX, ri, si, b, 0, lane)]) = ahi[b];
    *reinterpret_cast<vec*>(&g_mx_tab[mx_slot(MX, ri, si, b, 1, lane)]) = alo[b];
  }
}

// One workgroup of 64 threads fills every slot.
static __global__ __launch_bounds__(64) void mx_table_init() {
  const int lane = (int)threadIdx.x;
  mx_store<1, 16, -1>(lane);
  mx_store<1, 16, 1>(lane);
  mx_store<1, 8, -1>(lane);
  mx_store<1, 8, 1>(lane);
  mx_store<2, 16, -1>(lane);
  mx_store<2, 16, 1>(lane);
  mx_store<2, 8, -1>(lane);
  mx_store<2, 8, 1>(lane);
}

// One stage (radix RAD = 8 or 16, Ns = product of the earlier radices) over
// all C line slots of the workgroup's LDS tile; every thread of the
// workgroup (whole waves) takes part.  Same data movement as stage_lds:
// butterfly jb of a line reads k = jb + r*(n/RAD), twiddles by
// W_{Ns RAD}^{(jb%Ns) r}, writes k = (jb/Ns)*Ns*RAD + jb%Ns + r*Ns.
template <int RAD, int SIGN, int MX>
__device__ __forceinline__ void stage_mfma(cpx<float>* lds, int C, int stride, int n, int Ns,
                                           sa_fastdiv_t nsd, const cpx<float>* __restrict__ tw) {
  static_assert(RAD == 8 || RAD == 16, "matrix-core stage covers radix 8 and 16");
  constexpr int KPL = RAD == 16 ? 8 : 4;  // K values per lane
  constexpr int IN = KPL / 2;             // complex inputs per lane
  constexpr int MB = RAD / 8;             // 16-row output blocks
  constexpr int MAXR = 64 / RAD;          // rounds: C*n/RAD <= 16 * blockDim / RAD
  typedef MxVec<MX, KPL> V;
  typedef typename V::type vec;
  constexpr float LO = MX == 2 ? 2048.f : 1.f;  // scale of the lo halves
  constexpr float ILO = 1.f / LO;

  const int lane = (int)(threadIdx.x & 63u);
  const int w = (int)(threadIdx.x >> 6);
  const int nw = (int)(blockDim.x >> 6);
  const int j = lane & 15, h = lane >> 4;
  const int nb = n / RAD;
  const int nbt = C * nb;
  const FastDiv fd{nsd.d, nsd.m, nsd.s};

  // Constant DFT operand (rows 16b + (lane & 15), columns KPL*h + e), from
  // the table mx_table_init filled.
  vec ahi[MB], alo[MB];
  {
    const int ri = RAD == 16 ? 1 : 0, si = SIGN < 0 ? 0 : 1;
#pragma unroll
    for (int b = 0; b < MB; ++b) {
      ahi[b] = *reinterpret_cast<const vec*>(&g_mx_tab[mx_slot(MX, ri, si, b, 0, lane)]);
      alo[b] = *reinterpret_cast<const vec*>(&g_mx_tab[mx_slot(MX, ri, si, b, 1, lane)]);
    }
  }

  cpx<float> res[MAXR][MB][2];
#pragma unroll
  for (int r = 0; r < MAXR; ++r) {
    const int g = (r * nw + w) * 16 + j;
    const bool valid = g < nbt;
    const int c = valid ? g / nb : 0;
    const int jb = valid ? g - c * nb : 0;
    const cpx<float>* line = lds + (size_t)c * stride;
    float xs[KPL];
    int rem = 0;
    if (Ns > 1) rem = jb - (int)fdiv((uint32_t)jb, fd) * Ns;
#pragma unroll
    for (int e = 0; e < IN; ++e) {
      const int m = IN * h + e;
      cpx<float> x = {0.f, 0.f};
      if (valid) {
        x = line[lp(jb + m * nb)];
        if (Ns > 1 && m > 0) x = cmul_dir<SIGN>(x, tw[rem * (RAD - 1) + m - 1]);
      }
      xs[2 * e] = x.x;
      xs[2 * e + 1] = x.y;
    }
    float sc = 1.f, isc = 1.f;
    if constexpr (MX == 2) {
      float mx = 0.f;
#pragma unroll
      for (int e = 0; e < KPL; ++e) mx = fmaxf(mx, fabsf(xs[e]));
      mx = fmaxf(mx, __shfl_xor(mx, 16));
      mx = fmaxf(mx, __shfl_xor(mx, 32));
      int ex = 0;
      (void)frexpf(mx, &ex);
      sc = ldexpf(1.f, -ex);
      isc = ldexpf(1.f, ex);
    }
    vec xhi, xlo;
#pragma unroll
    for (int e = 0; e < KPL; ++e) {
      const float x = xs[e] * sc;
      V::set(xhi, e, x);
      V::set(xlo, e, (x - V::get(xhi, e)) * LO);
    }
#pragma unroll
    for (int b = 0; b < MB; ++b) {
      const mx_f32x4 z = {0.f, 0.f, 0.f, 0.f};
      const mx_f32x4 big = mx_mma<MX, KPL>(ahi[b], xhi, z);
      mx_f32x4 small = mx_mma<MX, KPL>(ahi[b], xlo, z);
      small = mx_mma<MX, KPL>(alo[b], xhi, small);
      // rows 4h..4h+3 of block b: outputs p = 8b + 2h (Re, Im), p + 1 (Re, Im)
      res[r][b][0] = {(big[0] + small[0] * ILO) * isc, (big[1] + small[1] * ILO) * isc};
      res[r][b][1] = {(big[2] + small[2] * ILO) * isc, (big[3] + small[3] * ILO) * isc};
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < MAXR; ++r) {
    const int g = (r * nw + w) * 16 + j;
    if (g < nbt) {
      const int c = g / nb;
      const int jb = g - c * nb;
      const int q = (int)fdiv((uint32_t)jb, fd);
      const int rem = jb - q * Ns;
      const int idx = q * Ns * RAD + rem;
      cpx<float>* line = lds + (size_t)c * stride;
#pragma unroll
      for (int b = 0; b < MB; ++b) {
        const int p = 8 * b + 2 * h;
        line[lp(idx + p * Ns)] = res[r][b][0];
        line[lp(idx + (p + 1) * Ns)] = res[r][b][1];
      }
    }
  }
  __syncthreads();
}

}  // namespace sa
