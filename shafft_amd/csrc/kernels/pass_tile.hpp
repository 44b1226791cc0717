// Device code of one LDS-resident FFT pass tile (C lines), shared by the
// per-pass launch (fft_pass.hip) and the persistent batch launch
// (fft_batch.hip).  See fft_pass.hip for the scheme.
#pragma once

#include "fft_device.hpp"
#include "pass_desc.h"

namespace sa {

// Register budget per thread per stage, in complex elements.
constexpr int EMAX = 16;

// LDS index of element k of a line.  Padded rows (PAD): one pad element
// every 16 (k -> k + k/16) keeps the stride-R butterfly writes of the early
// stages off a single bank; the empty asm keeps the sum whole, so the
// address is one shift-and-add from it (the compiler otherwise shifts both
// terms apart and adds three ways: one more VALU instruction per access).
// Unpadded rows (!PAD) are for passes whose stage accesses already spread
// over the banks (odd radices, line-fast lanes): an access is then the
// line's base plus a wave-uniform offset, one add instead of four VALU
// instructions, and the register stages' accesses are immediate offsets.
template <bool PAD = true>
__device__ __forceinline__ int lpx(int k) {
  if constexpr (PAD) {
    int v = k + (k >> 4);
    asm("" : "+v"(v));
    return v;
  } else {
    return k;
  }
}
__device__ __forceinline__ int lp(int k) { return lpx<true>(k); }

// Wave-uniform copy of a value that is the same in every lane (descriptor
// fields read from memory).  Control flow that encloses a barrier must be
// provably uniform, or the compiler may give lanes of one wave different
// loop exits and the wave reaches the barriers out of step.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ uint32_t fdiv_d(uint32_t x, sa_fastdiv_t f) {
  return fdiv(x, FastDiv{f.d, f.m, f.s});
}

template <typename R>
struct Vec2;
template <>
struct Vec2<float> {
  typedef float type __attribute__((ext_vector_type(2)));
};
template <>
struct Vec2<double> {
  typedef double type __attribute__((ext_vector_type(2)));
};

template <typename R>
__device__ __forceinline__ cpx<R> gload(const cpx<R>* p, bool nt) {
  typedef typename Vec2<R>::type V;
  V v = nt ? __builtin_nontemporal_load(reinterpret_cast<const V*>(p)) : *reinterpret_cast<const V*>(p);
  return {v.x, v.y};
}

template <typename R>
__device__ __forceinline__ void gstore(cpx<R>* p, cpx<R> x, bool nt) {
  typedef typename Vec2<R>::type V;
  V v;
  v.x = x.x;
  v.y = x.y;
  if (nt)
    __builtin_nontemporal_store(v, reinterpret_cast<V*>(p));
  else
    *reinterpret_cast<V*>(p) = v;
}

// ---------------------------------------------------------------------------
// Stockham stages.  Stage s (radix RAD, Ns = product of earlier radices):
// butterfly jb reads k = jb + r*(n/RAD), twiddles by W_{Ns RAD}^{(jb%Ns) r},
// and writes k = (jb/Ns)*Ns*RAD + jb%Ns + r*Ns.
// ---------------------------------------------------------------------------

// LDS -> LDS, butterflies jb = tl + i*T held in registers across a barrier.
// A slot runs when any lane of the wave has a butterfly there (a wave-uniform
// branch); in a running slot a lane past the line's last butterfly
// (jb >= nb) computes the last one again and only its writes are skipped, so
// the reads, twiddles and DFT run without exec-mask branches (a guarded
// butterfly cost those branches and register moves at every join).  Ns is
// wave-uniform (the twiddle branch is a scalar branch).
template <typename R, int RAD, int SIGN, int E = EMAX, bool PAD = true>
__device__ __forceinline__ void stage_lds(cpx<R>* line, int n, int Ns, sa_fastdiv_t nsd,
                                          const cpx<R>* __restrict__ tw, int tl, int T) {
  constexpr int MAXB = (E + RAD - 1) / RAD;
  const int nb = n / RAD;
  const FastDiv fd{nsd.d, nsd.m, nsd.s};
  cpx<R> v[MAXB][RAD];
  int idx[MAXB];
  bool run[MAXB];
#pragma unroll
  for (int i = 0; i < MAXB; ++i) {
    run[i] = __any(tl + i * T < nb) != 0;
    if (!run[i]) continue;
    const int jb = min(tl + i * T, nb - 1);
#pragma unroll
    for (int r = 0; r < RAD; ++r) v[i][r] = line[lpx<PAD>(jb + r * nb)];
    const int q = (int)fdiv((uint32_t)jb, fd);
    const int rem = jb - q * Ns;
    if (Ns > 1) {
      const cpx<R>* w = tw + rem * (RAD - 1);
#pragma unroll
      for (int r = 1; r < RAD; ++r) v[i][r] = cmul_dir<SIGN>(v[i][r], w[r - 1]);
    }
    dft<RAD, SIGN>(v[i]);
    idx[i] = q * Ns * RAD + rem;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < MAXB; ++i) {
    if (run[i] && tl + i * T < nb) {
#pragma unroll
      for (int r = 0; r < RAD; ++r) line[lpx<PAD>(idx[i] + r * Ns)] = v[i][r];
    }
  }
  __syncthreads();
}

// Radices above 16: a line has fewer butterflies (n/RAD) than threads
// (n/16), so the tile's C*n/RAD butterflies are numbered densely over all
// C*T threads (butterfly g = t: line g / (n/RAD)), one round, rather than
// leaving a 1 - 16/RAD share of every line's threads idle.
template <typename R, int RAD, int SIGN, bool PAD = true>
__device__ __forceinline__ void stage_lds_wide(cpx<R>* lds, int stride, int C, int nthr, int n,
                                               int Ns, sa_fastdiv_t nsd, sa_fastdiv_t nbd,
                                               const cpx<R>* __restrict__ tw, int t) {
  static_assert(RAD > 16, "one butterfly per thread");
  const int nb = n / RAD;
  const FastDiv fd{nsd.d, nsd.m, nsd.s};
  const FastDiv fb{nbd.d, nbd.m, nbd.s};
  const bool ok = t < nthr && t < C * nb;
  const int c = ok ? (int)fdiv((uint32_t)t, fb) : 0;
  const int jb = t - c * nb;
  cpx<R>* line = lds + (size_t)c * stride;
  cpx<R> v[RAD];
  if (ok) {
#pragma unroll
    for (int r = 0; r < RAD; ++r) v[r] = line[lpx<PAD>(jb + r * nb)];
    if (Ns > 1) {
      const int q = (int)fdiv((uint32_t)jb, fd);
      const int rem = jb - q * Ns;
      const cpx<R>* w = tw + rem * (RAD - 1);
#pragma unroll
      for (int r = 1; r < RAD; ++r) v[r] = cmul_dir<SIGN>(v[r], w[r - 1]);
    }
    dft<RAD, SIGN>(v);
  }
  __syncthreads();
  if (ok) {
    const int q = (int)fdiv((uint32_t)jb, fd);
    const int rem = jb - q * Ns;
    const int idx = q * Ns * RAD + rem;
#pragma unroll
    for (int r = 0; r < RAD; ++r) line[lpx<PAD>(idx + r * Ns)] = v[r];
  }
  __syncthreads();
}

}  // namespace sa
#include "mfma_stage.hpp"
namespace sa {

// Stage 0 from registers (v[m] = element tl + T*m, n == EMAX*T) into LDS.
template <typename R, int RAD, int SIGN, bool PAD = true>
__device__ __forceinline__ void stage_first_regs(const cpx<R> (&v)[EMAX], cpx<R>* line, int tl,
                                                 int T, bool act) {
  constexpr int MAXB = EMAX / RAD;
#pragma unroll
  for (int i = 0; i < MAXB && act; ++i) {
    cpx<R> u[RAD];
#pragma unroll
    for (int r = 0; r < RAD; ++r) u[r] = v[i + r * MAXB];
    dft<RAD, SIGN>(u);
    const int idx = (tl + i * T) * RAD;
#pragma unroll
    for (int r = 0; r < RAD; ++r) line[lpx<PAD>(idx + r)] = u[r];
  }
  __syncthreads();
}

// Last stage (Ns = n/RAD) from LDS into registers: v[i + r*MAXB] receives
// output element jb + r*Ns = tl + T*(i + r*MAXB).
template <typename R, int RAD, int SIGN, bool PAD = true>
__device__ __forceinline__ void stage_last_regs(cpx<R> (&v)[EMAX], const cpx<R>* line, int n,
                                                const cpx<R>* __restrict__ tw, int tl, int T,
                                                bool act) {
  constexpr int MAXB = EMAX / RAD;
  const int nb = n / RAD;
#pragma unroll
  for (int i = 0; i < MAXB && act; ++i) {
    const int jb = tl + i * T;
    cpx<R> u[RAD];
#pragma unroll
    for (int r = 0; r < RAD; ++r) u[r] = line[lpx<PAD>(jb + r * nb)];
    const cpx<R>* w = tw + jb * (RAD - 1);
#pragma unroll
    for (int r = 1; r < RAD; ++r) u[r] = cmul_dir<SIGN>(u[r], w[r - 1]);
    dft<RAD, SIGN>(u);
#pragma unroll
    for (int r = 0; r < RAD; ++r) v[i + r * MAXB] = u[r];
  }
}

// The butterfly lane index as the stage loop sees it: an empty asm statement
// marks it as redefined every stage, so the compiler computes each stage's
// LDS addresses inside that stage's radix case instead of hoisting the
// address sets of every radix case out of the stage loop (all of them live
// at once: +80 VGPRs, and computed per tile for cases that never run).
__device__ __forceinline__ int stage_lane(uint32_t tl) {
  int v = (int)tl;
  asm volatile("" : "+v"(v));
  return v;
}

// One Stockham stage of radix `rad` (wave-uniform) on this thread's line.
template <typename R, int SIGN, int E = EMAX, bool PAD = true>
__device__ __forceinline__ void stage_run(int rad, int n, int Ns, sa_fastdiv_t nsd, sa_fastdiv_t nbd,
                                          const cpx<R>* tws, cpx<R>* line, int tl, int T,
                                          cpx<R>* lds, int stride, int C, int t) {
  switch (rad) {
    case 2: stage_lds<R, 2, SIGN, E, PAD>(line, n, Ns, nsd, tws, tl, T); break;
    case 3: stage_lds<R, 3, SIGN, E, PAD>(line, n, Ns, nsd, tws, tl, T); break;
    case 4: stage_lds<R, 4, SIGN, E, PAD>(line, n, Ns, nsd, tws, tl, T); break;
    case 5: stage_lds<R, 5, SIGN, E, PAD>(line, n, Ns, nsd, tws, tl, T); break;
    case 6: stage_lds<R, 6, SIGN, E, PAD>(line, n, Ns, nsd, tws, tl, T); break;
    case 7: stage_lds<R, 7, SIGN, E, PAD>(line, n, Ns, nsd, tws, tl, T); break;
    case 8: stage_lds<R, 8, SIGN, E, PAD>(line, n, Ns, nsd, tws, tl, T); break;
    case 9: stage_lds<R, 9, SIGN, E, PAD>(line, n, Ns, nsd, tws, tl, T); break;
    case 10: stage_lds<R, 10, SIGN, E, PAD>(line, n, Ns, nsd, tws, tl, T); break;
    case 12: stage_lds<R, 12, SIGN, E, PAD>(line, n, Ns, nsd, tws, tl, T); break;
    case 16: stage_lds<R, 16, SIGN, E, PAD>(line, n, Ns, nsd, tws, tl, T); break;
    case 14: stage_lds<R, 14, SIGN, E, PAD>(line, n, Ns, nsd, tws, tl, T); break;
    case 15: stage_lds<R, 15, SIGN, E, PAD>(line, n, Ns, nsd, tws, tl, T); break;
#define SA_WIDE(RR) \
    case RR: stage_lds_wide<R, RR, SIGN, PAD>(lds, stride, C, C * T, n, Ns, nsd, nbd, tws, t); break;
    SA_WIDE(18) SA_WIDE(20) SA_WIDE(21)
#undef SA_WIDE
    // Radices 24, 25, 28 for complex double only: in the complex-float
    // kernel they would raise the register count past 5 waves per SIMD.
#define SA_WIDE_D(RR)                                                                    \
    case RR:                                                                             \
      if constexpr (sizeof(R) == 8)                                                      \
        stage_lds_wide<R, RR, SIGN, PAD>(lds, stride, C, C * T, n, Ns, nsd, nbd, tws, t);     \
      break;
    SA_WIDE_D(24) SA_WIDE_D(25) SA_WIDE_D(28)
#undef SA_WIDE_D
    default: break;
  }
}

// MX != 0 (complex float only): radix-8/16 stages on the matrix cores in
// split precision (mfma_stage.hpp); `lds` is the tile's LDS base.
// twb: the stage twiddle table (global, or staged in LDS by the caller).
template <typename R, int SIGN, int MX = 0, int E = EMAX, bool PAD = true>
__device__ __forceinline__ void stage_dispatch(const sa_pass_t& a, int s, cpx<R>* line, int tl,
                                               cpx<R>* lds, const cpx<R>* twb) {
  const cpx<R>* tws = twb + a.twoff[s];
  if constexpr (MX != 0) {
    static_assert(sizeof(R) == 4 && PAD, "split-precision stages: complex float, padded rows");
    const int rad = uni(a.radix[s]);
    if (rad == 16) {
      stage_mfma<16, SIGN, MX>(lds, a.C, a.ldsstride, a.n, a.ns[s], a.nsdiv[s], tws);
      return;
    }
    if (rad == 8) {
      stage_mfma<8, SIGN, MX>(lds, a.C, a.ldsstride, a.n, a.ns[s], a.nsdiv[s], tws);
      return;
    }
  }
  stage_run<R, SIGN, E, PAD>(uni(a.radix[s]), a.n, uni(a.ns[s]), a.nsdiv[s], a.nbdiv[s], tws, line, tl, a.T,
                     lds, a.ldsstride, a.C, stage_lane(threadIdx.x));
}

template <typename R, int SIGN, bool PAD = true>
__device__ __forceinline__ void first_dispatch(const sa_pass_t& a, const cpx<R> (&v)[EMAX],
                                               cpx<R>* line, int tl, bool act) {
  switch (uni(a.radix[0])) {
    case 2: stage_first_regs<R, 2, SIGN, PAD>(v, line, tl, a.T, act); break;
    case 4: stage_first_regs<R, 4, SIGN, PAD>(v, line, tl, a.T, act); break;
    case 8: stage_first_regs<R, 8, SIGN, PAD>(v, line, tl, a.T, act); break;
    case 16: stage_first_regs<R, 16, SIGN, PAD>(v, line, tl, a.T, act); break;
    default: __syncthreads(); break;
  }
}

template <typename R, int SIGN, bool PAD = true>
__device__ __forceinline__ void last_dispatch(const sa_pass_t& a, cpx<R> (&v)[EMAX],
                                              const cpx<R>* line, int tl, bool act,
                                              const cpx<R>* twb) {
  const int s = a.nstages - 1;
  const cpx<R>* tws = twb + a.twoff[s];
  switch (a.radix[s]) {
    case 2: stage_last_regs<R, 2, SIGN, PAD>(v, line, a.n, tws, tl, a.T, act); break;
    case 4: stage_last_regs<R, 4, SIGN, PAD>(v, line, a.n, tws, tl, a.T, act); break;
    case 8: stage_last_regs<R, 8, SIGN, PAD>(v, line, a.n, tws, tl, a.T, act); break;
    case 16: stage_last_regs<R, 16, SIGN, PAD>(v, line, a.n, tws, tl, a.T, act); break;
    default:  // not reached: register stages need a radix dividing 16 at both ends
#pragma unroll
      for (int i = 0; i < EMAX; ++i) v[i] = {R(0), R(0)};
      break;
  }
}

// ---------------------------------------------------------------------------
// Line addressing
// ---------------------------------------------------------------------------
struct LineOff {
  int64_t in, out;
  int64_t ld_base, st_base;  // digit terms of the load/store hooks
  uint64_t ptw_dig;          // digit driving the four-step twiddle
};

// digit[lvl] as arithmetic (no indexed access, which would spill to scratch)
__device__ __forceinline__ uint64_t sel3(int lvl, uint32_t d0, uint32_t d1, uint32_t d2) {
  return (uint64_t)((lvl == 0) ? d0 : 0u) + (uint64_t)((lvl == 1) ? d1 : 0u) +
         (uint64_t)((lvl == 2) ? d2 : 0u);
}

__device__ __forceinline__ LineOff line_offsets(const sa_pass_t& a, uint32_t b) {
  LineOff o;
  const uint32_t q2 = fdiv_d(b, a.n2div);
  const uint32_t d2 = b - q2 * (uint32_t)a.n2;
  const uint32_t d0 = fdiv_d(q2, a.n1div);
  const uint32_t d1 = q2 - d0 * (uint32_t)a.n1;
  o.in = (int64_t)d0 * a.is0 + (int64_t)d1 * a.is1 + (int64_t)d2 * a.is2;
  o.out = (int64_t)d0 * a.os0 + (int64_t)d1 * a.os1 + (int64_t)d2 * a.os2;
  o.ld_base = (int64_t)sel3(a.ld_level, d0, d1, d2) * a.ld_dmul +
              (int64_t)sel3(a.ld_level2, d0, d1, d2) * a.ld_dmul2;
  o.st_base = (int64_t)sel3(a.st_level, d0, d1, d2) * a.st_dmul +
              (int64_t)sel3(a.st_level2, d0, d1, d2) * a.st_dmul2;
  o.ptw_dig = sel3(a.ptw_level, d0, d1, d2);
  return o;
}

// Line slot of thread t for a mapping kind: k-fast t = c*T + k0,
// line-fast t = k0*C + c.
__device__ __forceinline__ uint32_t slot_line(const sa_pass_t& a, uint32_t t, bool kfast) {
  if (kfast) return fdiv_d(t, a.Tdiv);
  return t - fdiv_d(t, a.Cdiv) * (uint32_t)a.C;
}
__device__ __forceinline__ uint32_t slot_elem(const sa_pass_t& a, uint32_t t, bool kfast) {
  if (kfast) return t - fdiv_d(t, a.Tdiv) * (uint32_t)a.T;
  return fdiv_d(t, a.Cdiv);
}

// Elements per group of the hooked loads and stores: the group's data and
// table values are held in registers at once.
template <typename R, int E>
__device__ constexpr int hook_group() {
  return E < 8 ? E : 8;
}

// Loads elements k = k0 + i*T (i < E, k < n) of one line into v, with the
// load hooks applied (zero outside the hook's limit).  FULL: n == E*T and
// k0 < T, so every k is in the line and no element load is predicated (a
// predicated load into the register array that the first stage then reads
// makes the compiler park elements in scratch and wait on each load).
template <typename R, int E = EMAX, bool FULL = false>
__device__ __forceinline__ void load_line(const sa_pass_t& a, const LineOff& o, uint32_t k0,
                                          cpx<R> (&v)[E]) {
  const int n = a.n;
  const uint32_t T = (uint32_t)a.T;
  const cpx<R>* __restrict__ src = reinterpret_cast<const cpx<R>*>(a.in) + o.in;
  const bool nt = a.nt != 0;
  const bool plain = a.pre == nullptr && a.ld_level < 0 && a.ld_kmul == 1 && a.ld_limit >= n;
  if (plain) {
    // Slots past the line's end load its last element (same address as a
    // live lane, no extra traffic) instead of being predicated; the caller
    // drops them.
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const uint32_t k = FULL ? k0 + i * T : min(k0 + i * T, (uint32_t)n - 1u);
      v[i] = gload(src + (int64_t)k * a.iks, nt);
    }
  } else {
    // Hooked elements in groups of HG: every data and table load of a group
    // is issued before any is used.  (A table multiply inside each
    // element's own conditional waits for that element's two loads, so the
    // loads of a thread ran one after another at full memory latency.)
    const cpx<R>* __restrict__ pre = reinterpret_cast<const cpx<R>*>(a.pre);
    const int64_t base = o.ld_base;
    constexpr int HG = hook_group<R, E>();
#pragma unroll
    for (int g = 0; g < E; g += HG) {
      cpx<R> x[HG], w[HG];
#pragma unroll
      for (int j = 0; j < HG; ++j) {
        const uint32_t k = k0 + (g + j) * T;
        const int64_t m = (int64_t)k * a.ld_kmul + base;
        const bool ok = (FULL || k < (uint32_t)n) && m < a.ld_limit;
        x[j] = {R(0), R(0)};
        w[j] = {R(1), R(0)};
        if (ok) x[j] = gload(src + (int64_t)k * a.iks, nt);
        if (ok && pre) w[j] = pre[m];
      }
#pragma unroll
      for (int j = 0; j < HG; ++j) {
        const uint32_t k = k0 + (g + j) * T;
        if (FULL || k < (uint32_t)n) v[g + j] = pre ? cmul(x[j], w[j]) : x[j];
      }
    }
  }
}

// Stores elements k = k0 + i*T of one line from v with the store hooks
// (FULL as for load_line).
template <typename R, int SIGN, int E = EMAX, bool FULL = false>
__device__ __forceinline__ void store_line(const sa_pass_t& a, const LineOff& o, uint32_t k0,
                                           const cpx<R> (&v)[E]) {
  const int n = a.n;
  const uint32_t T = (uint32_t)a.T;
  cpx<R>* __restrict__ dst = reinterpret_cast<cpx<R>*>(a.out) + o.out;
  const cpx<R>* __restrict__ post = reinterpret_cast<const cpx<R>*>(a.post);
  const bool nt = a.nt != 0;
  if (a.ptw_level >= 0) {
    // W_N^{(digit * k) mod N} for k = k0 + i*T is geometric in i: look up
    // W^{e(k0)} and the step W^{digit*T} once (two-level table) and form the
    // rest by FP64 complex products (<= 15 deep), not two scattered table
    // loads per element.
    const uint64_t N = (uint64_t)a.ptw_N;
    const uint64_t dig = o.ptw_dig;
    const cpx<R>* __restrict__ lo = reinterpret_cast<const cpx<R>*>(a.ptw_lo);
    const cpx<R>* __restrict__ hi = reinterpret_cast<const cpx<R>*>(a.ptw_hi);
    const uint64_t mask = (1ull << a.ptw_shift) - 1;
    auto w64 = [&](uint64_t e) {
      const cpx<R> h = hi[e >> a.ptw_shift], l = lo[e & mask];
      return cmul(cpx<double>{(double)h.x, (double)h.y}, cpx<double>{(double)l.x, (double)l.y});
    };
    cpx<double> w = w64((dig * k0) % N);
    const cpx<double> step = w64((dig * T) % N);
    const R cs = a.ptw_conj ? R(-1) : R(1);  // conj(W) on request
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const uint32_t k = k0 + i * T;
      const cpx<R> x = cmul_dir<SIGN>(v[i], cpx<R>{(R)w.x, cs * (R)w.y});
      if (FULL || k < (uint32_t)n) gstore(dst + (int64_t)k * a.oks, x, nt);
      if (i + 1 < E) w = cmul(w, step);
    }
  } else if (post == nullptr && a.st_level < 0 && a.st_kmul == 1 && a.st_limit >= n) {
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const uint32_t k = k0 + i * T;
      if (FULL || k < (uint32_t)n) {
        cpx<R> x = v[i];
        gstore(dst + (int64_t)k * a.oks, x, nt);
      }
    }
  } else {
    // As in load_line: a group's table loads are all issued before the
    // first of its stores needs one.
    const int64_t base = o.st_base;
    constexpr int HG = hook_group<R, E>();
#pragma unroll
    for (int g = 0; g < E; g += HG) {
      cpx<R> w[HG];
#pragma unroll
      for (int j = 0; j < HG; ++j) {
        const uint32_t k = k0 + (g + j) * T;
        const int64_t m = (int64_t)k * a.st_kmul + base;
        w[j] = {R(1), R(0)};
        if ((FULL || k < (uint32_t)n) && m < a.st_limit && post) w[j] = post[m];
      }
#pragma unroll
      for (int j = 0; j < HG; ++j) {
        const uint32_t k = k0 + (g + j) * T;
        const int64_t m = (int64_t)k * a.st_kmul + base;
        if ((FULL || k < (uint32_t)n) && m < a.st_limit)
          gstore(dst + (int64_t)k * a.oks, post ? cmul(v[g + j], w[j]) : v[g + j], nt);
      }
    }
  }
}

// Kernel variants: the generic LDS path, the register-resident first/last
// stage path (both flags set, same lane mapping on both sides), Bluestein.
enum PassMode { MODE_LDS = 0, MODE_REGS = 1, MODE_BLUE = 2 };

// Copies the stage twiddle table (a.tw_lds elements) into LDS after the C
// line slots; the callers place it after the tile's HBM loads are issued, so
// its L2 latency overlaps theirs instead of returning in every stage.
template <typename R>
__device__ __forceinline__ const cpx<R>* stage_twiddles_to_lds(const sa_pass_t& a, cpx<R>* lds) {
  const cpx<R>* __restrict__ g = reinterpret_cast<const cpx<R>*>(a.tw);
  cpx<R>* ltw = lds + (size_t)a.C * a.ldsstride;
  copy_to_lds<R>(ltw, g, a.tw_lds);
  return ltw;  // read after the next barrier
}

// Bluestein filter of line b (per-line offset when blu_level >= 0).
template <typename R>
__device__ __forceinline__ const cpx<R>* filter_of(const sa_pass_t& a, uint32_t b, bool active) {
  const cpx<R>* __restrict__ h = reinterpret_cast<const cpx<R>*>(a.blu_h);
  if (a.blu_level >= 0 && active && b < (uint32_t)a.batch) {
    const uint32_t q2 = fdiv_d(b, a.n2div);
    const uint32_t d2 = b - q2 * (uint32_t)a.n2;
    const uint32_t d0 = fdiv_d(q2, a.n1div);
    const uint32_t d1 = q2 - d0 * (uint32_t)a.n1;
    h += (int64_t)sel3(a.blu_level, d0, d1, d2) * a.blu_hstride;
  }
  return h;
}

// One tile (C consecutive lines) of a pass, run by the whole workgroup.
// Threads t >= C*T of a larger workgroup idle through the barriers.
// TWL: the stages read their twiddles from LDS (a.tw_lds > 0).
template <typename R, int SIGN, int MODE, int MX = 0, bool TWL = false, int E = EMAX, bool PAD = true>
__device__ __forceinline__ void pass_tile(const sa_pass_t& a, uint32_t tile, cpx<R>* lds) {
  static_assert(E == EMAX || MODE != MODE_REGS, "register stages hold EMAX elements");
  static_assert(PAD || (MODE != MODE_BLUE && MX == 0), "unpadded rows: plain and register-stage passes");
  constexpr bool BLUE = MODE == MODE_BLUE;

  const uint32_t nthr = (uint32_t)(a.C * a.T);
  const uint32_t t = threadIdx.x < nthr ? threadIdx.x : 0u;
  const bool active = threadIdx.x < nthr;
  const uint32_t line0 = tile * (uint32_t)a.C;
  const int n = a.n;
  constexpr bool reg_first = MODE == MODE_REGS;
  constexpr bool reg_last = MODE == MODE_REGS;

  // Thread identity for the transform: (line slot c, butterfly lane tl).
  const bool id_kfast = reg_first ? (a.ld_kfast != 0) : (a.st_kfast != 0);
  const uint32_t c = slot_line(a, t, id_kfast);
  // Inactive threads get a lane index past every butterfly so the stage
  // loops skip them; they still take part in every barrier.
  const uint32_t tl = active ? slot_elem(a, t, id_kfast) : (uint32_t)n;
  cpx<R>* line = lds + (size_t)c * a.ldsstride;
  cpx<R> v[E];
  const cpx<R>* twb = reinterpret_cast<const cpx<R>*>(a.tw);

  // ---- load ----------------------------------------------------------------
  if constexpr (reg_first) {
    const uint32_t b = line0 + c;
    if (active && b < (uint32_t)a.batch) {
      const LineOff o = line_offsets(a, b);
      load_line<R, E, true>(a, o, tl, v);
    } else {
#pragma unroll
      for (int i = 0; i < E; ++i) v[i] = {R(0), R(0)};
    }
    if constexpr (TWL) twb = stage_twiddles_to_lds<R>(a, lds);
    first_dispatch<R, SIGN, PAD>(a, v, line, (int)tl, active);
  } else {
    const uint32_t lc = slot_line(a, t, a.ld_kfast != 0);
    const uint32_t k0 = slot_elem(a, t, a.ld_kfast != 0);
    const uint32_t b = line0 + lc;
    const bool have = active && b < (uint32_t)a.batch;
    cpx<R> u[E];
    if (have) {
      const LineOff o = line_offsets(a, b);
      load_line<R, E>(a, o, k0, u);
    }
    if constexpr (TWL) twb = stage_twiddles_to_lds<R>(a, lds);
    if (have) {
      cpx<R>* ll = lds + (size_t)lc * a.ldsstride;
#pragma unroll
      for (int i = 0; i < E; ++i) {
        const uint32_t k = k0 + i * (uint32_t)a.T;
        if (k < (uint32_t)n) ll[lpx<PAD>((int)k)] = u[i];
      }
    }
    __syncthreads();
  }

  // ---- stages ----------------------------------------------------------------
  if constexpr (!BLUE) {
    const int s0 = reg_first ? 1 : 0;
    const int s1 = reg_last ? uni(a.nstages) - 1 : uni(a.nstages);
    for (int s = s0; s < s1; ++s) stage_dispatch<R, SIGN, MX, E, PAD>(a, s, line, stage_lane(tl), lds, twb);
  } else {
    const int nst = uni(a.nstages);
    for (int s = 0; s < nst; ++s) stage_dispatch<R, -1, 0, E>(a, s, line, stage_lane(tl), lds, twb);
    const cpx<R>* __restrict__ h = filter_of<R>(a, line0 + c, active);
    for (int k = (int)tl; k < n; k += a.T) line[lp(k)] = cmul(line[lp(k)], h[k]);
    __syncthreads();
    for (int s = 0; s < nst; ++s) stage_dispatch<R, +1, 0, E>(a, s, line, stage_lane(tl), lds, twb);
  }

  // ---- store -----------------------------------------------------------------
  if constexpr (reg_last) {
    // The last stage's outputs go to their own array: the loaded values in v
    // are dead after the first stage and must not stay live across the
    // stages (they did when both shared v, and parts of v went to scratch).
    cpx<R> w[E];
    last_dispatch<R, SIGN, PAD>(a, w, line, (int)tl, active, twb);
    const uint32_t b = line0 + c;
    if (active && b < (uint32_t)a.batch) {
      const LineOff o = line_offsets(a, b);
      store_line<R, SIGN, E, true>(a, o, tl, w);
    }
  } else {
    const uint32_t sc = slot_line(a, t, a.st_kfast != 0);
    const uint32_t k0 = slot_elem(a, t, a.st_kfast != 0);
    const uint32_t b = line0 + sc;
    if (active && b < (uint32_t)a.batch) {
      const cpx<R>* ll = lds + (size_t)sc * a.ldsstride;
      cpx<R> u[E];
#pragma unroll
      for (int i = 0; i < E; ++i) u[i] = ll[lpx<PAD>((int)min(k0 + i * (uint32_t)a.T, (uint32_t)n - 1u))];
      const LineOff o = line_offsets(a, b);
      store_line<R, SIGN, E>(a, o, k0, u);
    }
  }
}

}  // namespace sa
