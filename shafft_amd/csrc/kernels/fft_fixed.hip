// Fixed-length persistent FFT pass for gfx950: power-of-two N known at
// compile time (32 .. 8192), the hot shapes of the slab path.
//
// Same pass semantics as fft_pass.hip (pass_desc.h addressing, four-step
// twiddle), specialised so that
//   * every Stockham index is a compile-time shift/or, and LDS addresses
//     are one base register plus immediate offsets;
//   * each thread owns E elements (k = tl + T*m, T = N/E threads per line;
//     E = 16, or 32 for 512/1024-point complex float); the first stage runs on
//     the loaded registers and, when load and store use the same lane
//     mapping, the last stage writes its butterfly outputs straight to HBM;
//   * the grid is persistent: each workgroup walks tiles b, b+G, b+2G, ...
//     and issues the HBM loads of its next tile before running the LDS
//     stages of the current one, so memory traffic overlaps the exchanges.
// One tile = C lines (C a power of two).  HBM-bound: 2 * N * sizeof(T)
// algorithmic bytes per line.
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

template <int N, int S>
struct StageInfo {
  static constexpr int R = Radices<N>::r[S];
  static constexpr int NS = S == 0 ? 1 : StageInfo<N, S - 1>::NS * StageInfo<N, S - 1>::R;
};
template <int N>
struct StageInfo<N, -1> {
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
template <typename R, int N>
__device__ __forceinline__ void load_tile(const sa_pass_t& a, uint32_t b, uint32_t tl,
                                          cpx<R> (&v)[Elems<R, N>::E]) {
  constexpr int E = Elems<R, N>::E, T = N / E;
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

template <typename R, int N, int MAXB>
__device__ __forceinline__ StoreCtx<R> store_ctx(const sa_pass_t& a, uint32_t b, uint32_t tl) {
  constexpr int T = N / Elems<R, N>::E;
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

// Store one output x of slot m with its four-step twiddle w.
template <typename R, int SIGN>
__device__ __forceinline__ void put(const sa_pass_t& a, const StoreCtx<R>& c, int m, cpx<double> w,
                                    cpx<R> x) {
  if (a.ptw_level >= 0) x = cmul_dir<SIGN>(x, cpx<R>{(R)w.x, (R)w.y});
  st(c.dst + m * c.step, x, a.nt != 0);
}

// Stage 0 from registers into LDS (Ns = 1: output index jb*RAD + r).
template <typename R, int N, int SIGN>
__device__ __forceinline__ void stage_first(const cpx<R> (&v)[Elems<R, N>::E], cpx<R>* line,
                                            uint32_t tl) {
  constexpr int E = Elems<R, N>::E, RAD = Radices<N>::r[0];
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
template <typename R, int N, int S, int SIGN>
__device__ __forceinline__ void stage_mid(cpx<R>* line, const cpx<R>* __restrict__ tw,
                                          uint32_t tl) {
  constexpr int E = Elems<R, N>::E;
  constexpr int RAD = StageInfo<N, S>::R, NS = StageInfo<N, S>::NS;
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
template <typename R, int N, int SIGN, int S>
__device__ __forceinline__ void mids(cpx<R>* line, const sa_pass_t& a, const cpx<R>* twb,
                                     uint32_t tl) {
  if constexpr (S < Radices<N>::n - 1) {
    __syncthreads();
    stage_mid<R, N, S, SIGN>(line, twb + a.twoff[S], tl);
    mids<R, N, SIGN, S + 1>(line, a, twb, tl);
  }
}

// Last stage (Ns = N/RAD): butterfly jb = tl + i*T yields output elements
// jb + r*Ns = tl + T*(i + r*MAXB).  STORE: write them straight to HBM;
// otherwise park them in LDS in natural order for a re-mapped store.
template <typename R, int N, int SIGN, bool STORE>
__device__ __forceinline__ void stage_last(const sa_pass_t& a, const cpx<R>* twb, cpx<R>* line,
                                           uint32_t tl, const StoreCtx<R>& c) {
  constexpr int E = Elems<R, N>::E, S = Radices<N>::n - 1;
  constexpr int RAD = StageInfo<N, S>::R;
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

// TWL: the stage twiddles (a.tw_lds elements) are copied into LDS after the
// C line slots once per workgroup, so the stages read them with LDS loads
// and the prefetched HBM loads of the next tile stay in flight through the
// stages (a global twiddle load would wait for them: one in-order counter).
template <typename R, int N, int SIGN, bool TWL>
__global__ __launch_bounds__(512) void fft_fixed_pass(const sa_pass_t a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  cpx<R>* lds = reinterpret_cast<cpx<R>*>(lds_raw);
  constexpr int E = Elems<R, N>::E, T = N / E;
  const uint32_t t = threadIdx.x;
  const int logC = a.logC;
  const int stride = a.ldsstride;
  const uint32_t C = 1u << logC;
  const cpx<R>* twb = reinterpret_cast<const cpx<R>*>(a.tw);
  if constexpr (TWL) {
    cpx<R>* ltw = lds + ((size_t)stride << logC);
    for (int i = (int)t; i < a.tw_lds; i += (int)blockDim.x) ltw[i] = twb[i];
    twb = ltw;  // visible to all threads after the first stage's barrier
  }
  const bool same = a.ld_kfast == a.st_kfast;
  uint32_t c, tl, sc, stl;
  ident<T>(t, a.ld_kfast != 0, logC, c, tl);
  ident<T>(t, a.st_kfast != 0, logC, sc, stl);
  cpx<R>* line = lds + (size_t)c * stride;

  const uint32_t tiles = (uint32_t)((a.batch + C - 1) >> logC);
  // Tile order: workgroup b walks b, b+G, ... (workgroups are dealt to the
  // 8 XCDs round robin, so each XCD sees every 8th tile), or with xcd_order
  // the workgroups of XCD x = b % 8 walk their own contiguous eighth.
  uint32_t tile = blockIdx.x, step = gridDim.x, end = tiles;
  if (a.xcd_order && (gridDim.x & 7u) == 0) {
    const uint32_t x = blockIdx.x & 7u, per = (tiles + 7u) >> 3;
    tile = x * per + (blockIdx.x >> 3);
    step = gridDim.x >> 3;
    end = min(tiles, (x + 1u) * per);
  }
  cpx<R> v[E];
  if (tile < end) load_tile<R, N>(a, (tile << logC) + c, tl, v);
  while (tile < end) {
    stage_first<R, N, SIGN>(v, line, tl);
    const uint32_t next = tile + step;
    if (next < end) load_tile<R, N>(a, (next << logC) + c, tl, v);  // prefetch
    mids<R, N, SIGN, 1>(line, a, twb, tl);
    __syncthreads();
    if (same) {
      const StoreCtx<R> ctx =
          store_ctx<R, N, E / StageInfo<N, Radices<N>::n - 1>::R>(a, (tile << logC) + c, tl);
      stage_last<R, N, SIGN, true>(a, twb, line, tl, ctx);
    } else {
      // Transposing pass: natural order through LDS, then the store mapping.
      const StoreCtx<R> dummy = {};
      stage_last<R, N, SIGN, false>(a, twb, line, tl, dummy);
      __syncthreads();
      const StoreCtx<R> ctx = store_ctx<R, N, 1>(a, (tile << logC) + sc, stl);
      if (ctx.ok) {
        const cpx<R>* sl = lds + (size_t)sc * stride;
        cpx<double> w = ctx.w0;
#pragma unroll
        for (int m = 0; m < E; ++m) {
          put<R, SIGN>(a, ctx, m, w, sl[lp((int)stl + T * m)]);
          if (m + 1 < E) w = cmul(w, ctx.s1);
        }
      }
    }
    __syncthreads();
    tile = next;
  }
}

template <typename R, int N, int SIGN, bool TWL>
hipError_t launch_nt(const sa_pass_t* p, hipStream_t s, int grid) {
  const uint32_t nthr = (uint32_t)((N / Elems<R, N>::E) << p->logC);
  const size_t lds = (((size_t)p->ldsstride << p->logC) + (TWL ? (size_t)p->tw_lds : 0)) * sizeof(cpx<R>);
  static const bool attr = [] {  // once per instantiation, thread-safe
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&fft_fixed_pass<R, N, SIGN, TWL>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    (void)hipGetLastError();
    return true;
  }();
  (void)attr;
  hipLaunchKernelGGL((fft_fixed_pass<R, N, SIGN, TWL>), dim3((uint32_t)grid), dim3(nthr), lds, s, *p);
  return hipGetLastError();
}

template <typename R, int N, int SIGN>
hipError_t launch_n(const sa_pass_t* p, hipStream_t s, int grid) {
  return p->tw_lds > 0 ? launch_nt<R, N, SIGN, true>(p, s, grid) : launch_nt<R, N, SIGN, false>(p, s, grid);
}

template <typename R, int SIGN>
hipError_t launch_r(const sa_pass_t* p, hipStream_t s, int grid) {
  switch (p->n) {
    case 32: return launch_n<R, 32, SIGN>(p, s, grid);
    case 64: return launch_n<R, 64, SIGN>(p, s, grid);
    case 128: return launch_n<R, 128, SIGN>(p, s, grid);
    case 256: return launch_n<R, 256, SIGN>(p, s, grid);
    case 512: return launch_n<R, 512, SIGN>(p, s, grid);
    case 1024: return launch_n<R, 1024, SIGN>(p, s, grid);
    case 2048: return launch_n<R, 2048, SIGN>(p, s, grid);
    case 4096: return launch_n<R, 4096, SIGN>(p, s, grid);
    case 8192: return launch_n<R, 8192, SIGN>(p, s, grid);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace fx
}  // namespace sa

extern "C" int sa_fixed_supported(int n) {
  return n == 32 || n == 64 || n == 128 || n == 256 || n == 512 || n == 1024 || n == 2048 ||
         n == 4096 || n == 8192;
}

extern "C" int sa_fixed_elems(int n, int prec) { return (prec == 0 && (n == 512 || n == 1024)) ? 32 : 16; }

extern "C" int sa_fixed_max_threads(int prec) { (void)prec; return 512; }

extern "C" int sa_launch_fixed(const sa_pass_t* p, int prec, int sign, int grid, void* stream) {
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (prec == 0)
    return sign < 0 ? (int)sa::fx::launch_r<float, -1>(p, s, grid)
                    : (int)sa::fx::launch_r<float, +1>(p, s, grid);
  return sign < 0 ? (int)sa::fx::launch_r<double, -1>(p, s, grid)
                  : (int)sa::fx::launch_r<double, +1>(p, s, grid);
}

// Resident workgroups per CU for the fixed kernel of (n, prec) with the
// given block size and LDS bytes.
extern "C" int sa_fixed_occupancy(int n, int prec, int sign, int nthr, size_t lds) {
  int blocks = 0;
  const void* f = nullptr;
  // Forward variant with global twiddles; the LDS-twiddle variants use the
  // same registers (lds includes the table when it is staged).
#define SA_FX_CASE(NN)                                                                        \
  case NN:                                                                                    \
    f = prec == 0 ? (sign < 0 ? reinterpret_cast<const void*>(&sa::fx::fft_fixed_pass<float, NN, -1, false>)  \
                              : reinterpret_cast<const void*>(&sa::fx::fft_fixed_pass<float, NN, 1, false>))   \
                  : (sign < 0 ? reinterpret_cast<const void*>(&sa::fx::fft_fixed_pass<double, NN, -1, false>) \
                              : reinterpret_cast<const void*>(&sa::fx::fft_fixed_pass<double, NN, 1, false>)); \
    break;
  switch (n) {
    SA_FX_CASE(32)
    SA_FX_CASE(64)
    SA_FX_CASE(128)
    SA_FX_CASE(256)
    SA_FX_CASE(512)
    SA_FX_CASE(1024)
    SA_FX_CASE(2048)
    SA_FX_CASE(4096)
    SA_FX_CASE(8192)
    default: return 0;
  }
#undef SA_FX_CASE
  (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    (void)hipGetLastError();
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, f, nthr, lds) != hipSuccess) return 0;
  return blocks;
}
