// Compile-time-length Stockham stages shared by the fixed-length persistent
// pass (fft_fixed.hip) and the three-pass Bluestein kernels (fft_blue.hip):
// radix plans for power-of-two N (32 .. 8192), E elements per thread
// (k = tl + T*m), LDS rows padded by one element every 16, the first stage
// on registers and the last stage to HBM or back to registers.
#pragma once

#include "fft_device.hpp"
#include "pass_desc.h"

namespace sa {
namespace fx {

constexpr int ilog2(int v) { return v <= 1 ? 0 : 1 + ilog2(v / 2); }

// Elements per thread (16; 32 for 512/1024-point complex float, whose 32-
// and 16-line tiles need it) and the workgroup size bound of 512
// threads (256 VGPRs per lane); a 16-line tile of a 512-point complex double
// or 1024-point complex float line is one workgroup.
template <typename R, int N>
struct Elems {
  static constexpr int E = (sizeof(R) == 4 && (N == 512 || N == 1024)) ? 32 : 16;
  static constexpr int MAXT = 512;
};

// Radix sequence for N (each radix divides 16).
template <int N>
struct Radices;
template <> struct Radices<32> { static constexpr int n = 2, r[4] = {8, 4, 0, 0}; };
template <> struct Radices<64> { static constexpr int n = 2, r[4] = {8, 8, 0, 0}; };
template <> struct Radices<128> { static constexpr int n = 2, r[4] = {16, 8, 0, 0}; };
template <> struct Radices<256> { static constexpr int n = 2, r[4] = {16, 16, 0, 0}; };
template <> struct Radices<512> { static constexpr int n = 3, r[4] = {8, 8, 8, 0}; };
template <> struct Radices<1024> { static constexpr int n = 3, r[4] = {16, 8, 8, 0}; };
template <> struct Radices<2048> { static constexpr int n = 3, r[4] = {16, 16, 8, 0}; };
template <> struct Radices<4096> { static constexpr int n = 3, r[4] = {16, 16, 16, 0}; };
template <> struct Radices<8192> { static constexpr int n = 4, r[4] = {16, 16, 8, 4}; };

// The plan a kernel with E elements per thread runs: the radices above for
// E >= 16; with 8 elements per thread (the Bluestein row passes' wider
// workgroups) every radix divides 8.
template <int N, int E>
struct Plan : Radices<N> {};
template <> struct Plan<1024, 8> { static constexpr int n = 4, r[4] = {8, 8, 4, 4}; };
template <> struct Plan<2048, 8> { static constexpr int n = 4, r[4] = {8, 8, 8, 4}; };
template <> struct Plan<4096, 8> { static constexpr int n = 4, r[4] = {8, 8, 8, 8}; };

template <int N, int S, int E = 16>
struct StageInfo {
  static constexpr int R = Plan<N, E>::r[S];
  static constexpr int NS = S == 0 ? 1 : StageInfo<N, S - 1, E>::NS * StageInfo<N, S - 1, E>::R;
};
template <int N, int E>
struct StageInfo<N, -1, E> {
  static constexpr int R = 1, NS = 1;
};

__device__ __forceinline__ int lp(int k) { return k + (k >> 4); }

template <typename R>
struct V2;
template <>
struct V2<float> {
  typedef float t __attribute__((ext_vector_type(2)));
};
template <>
struct V2<double> {
  typedef double t __attribute__((ext_vector_type(2)));
};

template <typename R>
__device__ __forceinline__ cpx<R> ld(const cpx<R>* p, bool nt) {
  typedef typename V2<R>::t V;
  V v = nt ? __builtin_nontemporal_load(reinterpret_cast<const V*>(p)) : *reinterpret_cast<const V*>(p);
  return {v.x, v.y};
}
template <typename R>
__device__ __forceinline__ void st(cpx<R>* p, cpx<R> x, bool nt) {
  typedef typename V2<R>::t V;
  V v;
  v.x = x.x;
  v.y = x.y;
  if (nt)
    __builtin_nontemporal_store(v, reinterpret_cast<V*>(p));
  else
    *reinterpret_cast<V*>(p) = v;
}

__device__ __forceinline__ uint32_t fdv(uint32_t x, sa_fastdiv_t f) {
  return fdiv(x, FastDiv{f.d, f.m, f.s});
}

struct Line {
  int64_t in, out;
  uint64_t dig;  // four-step twiddle digit
};

__device__ __forceinline__ Line line_of(const sa_pass_t& a, uint32_t b) {
  const uint32_t q2 = fdv(b, a.n2div);
  const uint32_t d2 = b - q2 * (uint32_t)a.n2;
  const uint32_t d0 = fdv(q2, a.n1div);
  const uint32_t d1 = q2 - d0 * (uint32_t)a.n1;
  Line l;
  l.in = (int64_t)d0 * a.is0 + (int64_t)d1 * a.is1 + (int64_t)d2 * a.is2;
  l.out = (int64_t)d0 * a.os0 + (int64_t)d1 * a.os1 + (int64_t)d2 * a.os2;
  l.dig = (uint64_t)((a.ptw_level == 0) ? d0 : 0u) + (uint64_t)((a.ptw_level == 1) ? d1 : 0u) +
          (uint64_t)((a.ptw_level == 2) ? d2 : 0u);
  return l;
}

// Thread identity (line slot, lane in line) for a mapping kind.
template <int T>
__device__ __forceinline__ void ident(uint32_t t, bool kfast, int logC, uint32_t& c, uint32_t& tl) {
  if (kfast) {
    c = t >> ilog2(T);
    tl = t & (T - 1);
  } else {
    c = t & ((1u << logC) - 1);
    tl = t >> logC;
  }
}

// v[m] = element tl + T*m of line b (zeros past the last line).
template <typename R, int N, int E = Elems<R, N>::E>
__device__ __forceinline__ void load_tile(const sa_pass_t& a, uint32_t b, uint32_t tl,
                                          cpx<R> (&v)[E]) {
  constexpr int T = N / E;
  if (b < (uint32_t)a.batch) {
    const Line l = line_of(a, b);
    const cpx<R>* src = reinterpret_cast<const cpx<R>*>(a.in) + l.in + (int64_t)tl * a.iks;
    const int64_t step = (int64_t)T * a.iks;
    const bool nt = a.nt != 0;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      v[m] = ld(src, nt);
      src += step;
    }
  } else {
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = {R(0), R(0)};
  }
}

// Per-line store context: destination, element step and the four-step
// twiddle of slot m (element tl + T*m), W_Nt^{e(m)} with
// e(m) = dig * (tl + T*m) mod Nt.  The twiddles of a line are a geometric
// sequence, so each line looks up W^{e(0)} and the steps W^{es},
// W^{MAXB*es} once in the two-level table and forms the rest by FP64
// complex products (at most 24 deep: ~3e-15 relative), instead of two
// scattered table loads per element.
template <typename R>
struct StoreCtx {
  cpx<R>* dst;
  int64_t step;
  cpx<double> w0, s1, sb;  // W^{e(0)}, W^{es}, W^{MAXB*es}
  bool ok;
};

// W^e from the two-level table, widened to FP64.
template <typename R>
__device__ __forceinline__ cpx<double> ptw64(const sa_pass_t& a, uint64_t e) {
  const cpx<R>* lo = reinterpret_cast<const cpx<R>*>(a.ptw_lo);
  const cpx<R>* hi = reinterpret_cast<const cpx<R>*>(a.ptw_hi);
  const cpx<R> h = hi[e >> a.ptw_shift], l = lo[e & ((1ull << a.ptw_shift) - 1)];
  return cmul(cpx<double>{(double)h.x, (double)h.y}, cpx<double>{(double)l.x, (double)l.y});
}

template <typename R, int N, int MAXB, int E = Elems<R, N>::E>
__device__ __forceinline__ StoreCtx<R> store_ctx(const sa_pass_t& a, uint32_t b, uint32_t tl) {
  constexpr int T = N / E;
  StoreCtx<R> c;
  c.ok = b < (uint32_t)a.batch;
  const Line l = line_of(a, c.ok ? b : 0u);
  c.dst = reinterpret_cast<cpx<R>*>(a.out) + l.out + (int64_t)tl * a.oks;
  c.step = (int64_t)T * a.oks;
  c.w0 = c.s1 = c.sb = {1.0, 0.0};
  if (a.ptw_level >= 0) {
    const uint64_t Nt = (uint64_t)a.ptw_N;
    const uint64_t es = (l.dig * T) % Nt;
    c.w0 = ptw64<R>(a, (l.dig * tl) % Nt);
    c.s1 = ptw64<R>(a, es);
    c.sb = ptw64<R>(a, (es * MAXB) % Nt);
  }
  return c;
}

// The six table values behind a line's store twiddles (W^{e(0)}, W^{es},
// W^{MAXB*es}), loaded ahead of their use so the lookups' latency overlaps
// other work; store_ctx_from() forms the twiddles from them.
template <typename R>
struct PtwRaw {
  cpx<R> h[3], l[3];
};

template <typename R, int N, int MAXB, int E = Elems<R, N>::E>
__device__ __forceinline__ PtwRaw<R> ptw_fetch(const sa_pass_t& a, uint32_t b, uint32_t tl) {
  constexpr int T = N / E;
  PtwRaw<R> q;
#pragma unroll
  for (int j = 0; j < 3; ++j) q.h[j] = q.l[j] = {R(1), R(0)};
  if (a.ptw_level >= 0) {
    const Line l = line_of(a, b < (uint32_t)a.batch ? b : 0u);
    const uint64_t Nt = (uint64_t)a.ptw_N;
    const uint64_t es = (l.dig * T) % Nt;
    const uint64_t e[3] = {(l.dig * tl) % Nt, es, (es * MAXB) % Nt};
    const cpx<R>* lo = reinterpret_cast<const cpx<R>*>(a.ptw_lo);
    const cpx<R>* hi = reinterpret_cast<const cpx<R>*>(a.ptw_hi);
    const uint64_t mask = (1ull << a.ptw_shift) - 1;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      q.h[j] = hi[e[j] >> a.ptw_shift];
      q.l[j] = lo[e[j] & mask];
    }
  }
  return q;
}

template <typename R, int N, int E = Elems<R, N>::E>
__device__ __forceinline__ StoreCtx<R> store_ctx_from(const sa_pass_t& a, uint32_t b, uint32_t tl,
                                                      const PtwRaw<R>& q) {
  constexpr int T = N / E;
  StoreCtx<R> c;
  c.ok = b < (uint32_t)a.batch;
  const Line l = line_of(a, c.ok ? b : 0u);
  c.dst = reinterpret_cast<cpx<R>*>(a.out) + l.out + (int64_t)tl * a.oks;
  c.step = (int64_t)T * a.oks;
  auto w = [&](int j) {
    return cmul(cpx<double>{(double)q.h[j].x, (double)q.h[j].y}, cpx<double>{(double)q.l[j].x, (double)q.l[j].y});
  };
  c.w0 = c.s1 = c.sb = {1.0, 0.0};
  if (a.ptw_level >= 0) {
    c.w0 = w(0);
    c.s1 = w(1);
    c.sb = w(2);
  }
  return c;
}

// Store one output x of slot m with its four-step twiddle w.
template <typename R, int SIGN>
__device__ __forceinline__ void put(const sa_pass_t& a, const StoreCtx<R>& c, int m, cpx<double> w,
                                    cpx<R> x) {
  if (a.ptw_level >= 0) x = cmul_dir<SIGN>(x, cpx<R>{(R)w.x, (R)w.y});
  st(c.dst + m * c.step, x, a.nt != 0);
}

// Stage 0 from registers into LDS (Ns = 1: output index jb*RAD + r).
template <typename R, int N, int SIGN, int E = Elems<R, N>::E>
__device__ __forceinline__ void stage_first(const cpx<R> (&v)[E], cpx<R>* line,
                                            uint32_t tl) {
  constexpr int RAD = Plan<N, E>::r[0];
  constexpr int T = N / E, MAXB = E / RAD;
#pragma unroll
  for (int i = 0; i < MAXB; ++i) {
    cpx<R> u[RAD];
#pragma unroll
    for (int r = 0; r < RAD; ++r) u[r] = v[i + r * MAXB];
    dft<RAD, SIGN>(u);
    const int idx = (int)(tl + i * T) * RAD;
#pragma unroll
    for (int r = 0; r < RAD; ++r) line[lp(idx | r)] = u[r];
  }
}

// Middle stage S: LDS -> LDS (a barrier precedes the call).
template <typename R, int N, int S, int SIGN, int E = Elems<R, N>::E>
__device__ __forceinline__ void stage_mid(cpx<R>* line, const cpx<R>* __restrict__ tw,
                                          uint32_t tl) {
  constexpr int RAD = StageInfo<N, S, E>::R, NS = StageInfo<N, S, E>::NS;
  constexpr int T = N / E, MAXB = E / RAD, NB = N / RAD;
  cpx<R> u[MAXB][RAD];
#pragma unroll
  for (int i = 0; i < MAXB; ++i) {
    const int jb = (int)tl + i * T;
#pragma unroll
    for (int r = 0; r < RAD; ++r) u[i][r] = line[lp(jb + r * NB)];
    const int rem = jb & (NS - 1);
    const cpx<R>* w = tw + rem * (RAD - 1);
#pragma unroll
    for (int r = 1; r < RAD; ++r) u[i][r] = cmul_dir<SIGN>(u[i][r], w[r - 1]);
    dft<RAD, SIGN>(u[i]);
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < MAXB; ++i) {
    const int jb = (int)tl + i * T;
    const int rem = jb & (NS - 1);
    const int base = ((jb >> ilog2(NS)) << ilog2(NS * RAD)) | rem;
#pragma unroll
    for (int r = 0; r < RAD; ++r) line[lp(base + r * NS)] = u[i][r];
  }
}

// twb: the stage twiddle table, in LDS or in global memory.
template <typename R, int N, int SIGN, int S, int E = Elems<R, N>::E>
__device__ __forceinline__ void mids(cpx<R>* line, const sa_pass_t& a, const cpx<R>* twb,
                                     uint32_t tl) {
  if constexpr (S < Plan<N, E>::n - 1) {
    __syncthreads();
    stage_mid<R, N, S, SIGN, E>(line, twb + a.twoff[S], tl);
    mids<R, N, SIGN, S + 1, E>(line, a, twb, tl);
  }
}

// Last stage (Ns = N/RAD): butterfly jb = tl + i*T yields output elements
// jb + r*Ns = tl + T*(i + r*MAXB).  STORE: write them straight to HBM;
// otherwise park them in LDS in natural order for a re-mapped store.
template <typename R, int N, int SIGN, bool STORE, int E = Elems<R, N>::E>
__device__ __forceinline__ void stage_last(const sa_pass_t& a, const cpx<R>* twb, cpx<R>* line,
                                           uint32_t tl, const StoreCtx<R>& c) {
  constexpr int S = Plan<N, E>::n - 1;
  constexpr int RAD = StageInfo<N, S, E>::R;
  constexpr int T = N / E, MAXB = E / RAD, NB = N / RAD;
  const cpx<R>* tw = twb + a.twoff[S];
  if constexpr (STORE) {
    if (!c.ok) return;
    cpx<double> wi = c.w0;  // twiddle of slot i (r = 0)
#pragma unroll
    for (int i = 0; i < MAXB; ++i) {
      const int jb = (int)tl + i * T;
      cpx<R> u[RAD];
#pragma unroll
      for (int r = 0; r < RAD; ++r) u[r] = line[lp(jb + r * NB)];
      const cpx<R>* w = tw + jb * (RAD - 1);
#pragma unroll
      for (int r = 1; r < RAD; ++r) u[r] = cmul_dir<SIGN>(u[r], w[r - 1]);
      dft<RAD, SIGN>(u);
      cpx<double> wr = wi;  // slot i + r*MAXB
#pragma unroll
      for (int r = 0; r < RAD; ++r) {
        put<R, SIGN>(a, c, i + r * MAXB, wr, u[r]);
        if (r + 1 < RAD) wr = cmul(wr, c.sb);
      }
      if (i + 1 < MAXB) wi = cmul(wi, c.s1);
    }
  } else {
    cpx<R> u[MAXB][RAD];
#pragma unroll
    for (int i = 0; i < MAXB; ++i) {
      const int jb = (int)tl + i * T;
#pragma unroll
      for (int r = 0; r < RAD; ++r) u[i][r] = line[lp(jb + r * NB)];
      const cpx<R>* w = tw + jb * (RAD - 1);
#pragma unroll
      for (int r = 1; r < RAD; ++r) u[i][r] = cmul_dir<SIGN>(u[i][r], w[r - 1]);
      dft<RAD, SIGN>(u[i]);
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < MAXB; ++i)
#pragma unroll
      for (int r = 0; r < RAD; ++r) line[lp((int)tl + T * (i + r * MAXB))] = u[i][r];
  }
}

// Last stage into registers: out[m] = output element tl + T*m (slot
// m = i + r*MAXB of butterfly jb = tl + i*T), for element-wise work after
// the transform (Bluestein filter, chirp) before the next step.
template <typename R, int N, int SIGN, int E = Elems<R, N>::E>
__device__ __forceinline__ void stage_last_regs(const sa_pass_t& a, const cpx<R>* twb,
                                                const cpx<R>* line, uint32_t tl,
                                                cpx<R> (&out)[E]) {
  constexpr int S = Plan<N, E>::n - 1;
  constexpr int RAD = StageInfo<N, S, E>::R;
  constexpr int T = N / E, MAXB = E / RAD, NB = N / RAD;
  const cpx<R>* tw = twb + a.twoff[S];
#pragma unroll
  for (int i = 0; i < MAXB; ++i) {
    const int jb = (int)tl + i * T;
    cpx<R> u[RAD];
#pragma unroll
    for (int r = 0; r < RAD; ++r) u[r] = line[lp(jb + r * NB)];
    const cpx<R>* w = tw + jb * (RAD - 1);
#pragma unroll
    for (int r = 1; r < RAD; ++r) u[r] = cmul_dir<SIGN>(u[r], w[r - 1]);
    dft<RAD, SIGN>(u);
#pragma unroll
    for (int r = 0; r < RAD; ++r) out[i + r * MAXB] = u[r];
  }
}

}  // namespace fx
}  // namespace sa
