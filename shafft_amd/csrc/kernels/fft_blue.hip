// Three-pass Bluestein convolution on compile-time power-of-two lengths
// (gfx950).  The planner (PassProgram::bluestein_three_pass) factors the
// padded length M = N1 * N2 and keeps the spectrum in transposed order:
//
//   A  blue_cols<HK_CHIRP_IN>:  N1-point forward DFTs along the columns
//      (element stride N2) of the chirped, zero-padded input, times the
//      four-step twiddle W_M^{n2 k1} on the store;
//   B  blue_rows:               per row k1, the N2-point forward DFT, times
//      the row's filter H[k1 + N1 k2], the N2-point inverse DFT, times
//      conj(W_M^{n2 k1}) on the store -- one read and one write of the row;
//   C  blue_cols<HK_CHIRP_OUT>: N1-point inverse DFTs along the columns,
//      times the output chirp, stored only for indices below n (truncation).
//
// The stages are the fixed-length kernel's (fixed_stages.hpp): compile-time
// radix plan, first stage on the loaded registers, LDS rows padded every 16
// elements.  Unlike the generic pass, every table value an element needs
// (chirp, filter) is loaded together with the data, before any is used, so
// a workgroup waits for HBM once per tile instead of once per element, and a
// column tile spans 256 bytes of each row it touches (32 complex-float /
// 16 complex-double adjacent columns).  One tile per workgroup: the passes
// are short (M = 2^21 is one tile per CU or less), so the grid is the tile
// count and the per-workgroup latency, not a persistent loop, sets the time.
//
// HBM bytes per pass (complex T): A reads n elements and n chirp values and
// writes M; B reads M and M filter values and writes M; C reads M and n chirp
// values and writes n.
#include "fixed_stages.hpp"

namespace sa {
namespace fx {

enum BlueHook { HK_CHIRP_IN = 1, HK_CHIRP_OUT = 2 };

// Digit term of an element hook: digit[level] * dmul + digit[level2] * dmul2.
__device__ __forceinline__ int64_t hook_base(const sa_pass_t& a, uint32_t b, int level,
                                             int64_t dmul, int level2, int64_t dmul2) {
  const uint32_t q2 = fdv(b, a.n2div);
  const uint32_t d2 = b - q2 * (uint32_t)a.n2;
  const uint32_t d0 = fdv(q2, a.n1div);
  const uint32_t d1 = q2 - d0 * (uint32_t)a.n1;
  auto sel = [&](int lv) -> int64_t {
    return (int64_t)((lv == 0) ? d0 : 0u) + (int64_t)((lv == 1) ? d1 : 0u) + (int64_t)((lv == 2) ? d2 : 0u);
  };
  return (level >= 0 ? sel(level) * dmul : 0) + (level2 >= 0 ? sel(level2) * dmul2 : 0);
}

// Stage twiddles into LDS behind the C line slots (TWL), else global.
template <typename R, bool TWL>
__device__ __forceinline__ const cpx<R>* twiddle_base(const sa_pass_t& a, cpx<R>* lds) {
  const cpx<R>* twb = reinterpret_cast<const cpx<R>*>(a.tw);
  if constexpr (TWL) {
    cpx<R>* ltw = lds + ((size_t)a.ldsstride << a.logC);
    copy_to_lds<R>(ltw, twb, a.tw_lds);
    return ltw;  // read after the first stage's barrier
  }
  return twb;
}

// Passes A and C.  Line-fast on both sides: lanes walk the C adjacent
// columns of a tile, so a wavefront access covers C * sizeof(cpx) contiguous
// bytes of each of 64 / C rows.
template <typename R, int N, int HK, bool TWL, int E>
__global__ __launch_bounds__(512) void blue_cols(const sa_pass_t a) {
  constexpr int SIGN = HK == HK_CHIRP_IN ? -1 : +1;
  constexpr int T = N / E;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  cpx<R>* lds = reinterpret_cast<cpx<R>*>(lds_raw);
  const int logC = a.logC;
  uint32_t c, tl;
  ident<T>(threadIdx.x, false, logC, c, tl);
  cpx<R>* line = lds + (size_t)c * a.ldsstride;
  const uint32_t b = (blockIdx.x << logC) + c;
  const bool have = b < (uint32_t)a.batch;
  const Line l = line_of(a, have ? b : 0u);
  const bool nt = a.nt != 0;

  cpx<R> v[E];
  if constexpr (HK == HK_CHIRP_IN) {
    // x[m] * chirp[idx(m)] where idx(m) = k * kmul + digits < limit, else 0
    // (zero padding past n).  Out-of-range elements load from the buffer
    // bases (always valid) and are discarded, so no load is predicated.
    const cpx<R>* src = reinterpret_cast<const cpx<R>*>(a.in) + l.in;
    const cpx<R>* pre = reinterpret_cast<const cpx<R>*>(a.pre);
    const int64_t base = hook_base(a, have ? b : 0u, a.ld_level, a.ld_dmul, a.ld_level2, a.ld_dmul2);
    cpx<R> w[E];
    bool ok[E];
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int64_t k = (int64_t)tl + (int64_t)T * m;
      const int64_t idx = k * a.ld_kmul + base;
      ok[m] = have && idx < a.ld_limit;
      v[m] = ld(ok[m] ? src + k * a.iks : reinterpret_cast<const cpx<R>*>(a.in), nt);
      w[m] = pre[ok[m] ? idx : 0];
    }
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = ok[m] ? cmul(v[m], w[m]) : cpx<R>{R(0), R(0)};
  } else {
    load_tile<R, N, E>(a, b, tl, v);
  }
  const cpx<R>* twb = twiddle_base<R, TWL>(a, lds);
  stage_first<R, N, SIGN, E>(v, line, tl);

  if constexpr (HK == HK_CHIRP_IN) {
    // The four-step twiddle's table lookups go out before the stages.
    const PtwRaw<R> q = ptw_fetch<R, N, E / StageInfo<N, Plan<N, E>::n - 1, E>::R, E>(a, b, tl);
    mids<R, N, SIGN, 1, E>(line, a, twb, tl);
    __syncthreads();
    stage_last<R, N, SIGN, true, E>(a, twb, line, tl, store_ctx_from<R, N, E>(a, b, tl, q));
  } else {
    // Output chirp values of this thread's elements, loaded before the
    // stages so their latency hides behind the LDS exchanges.
    const cpx<R>* post = reinterpret_cast<const cpx<R>*>(a.post);
    const int64_t base = hook_base(a, have ? b : 0u, a.st_level, a.st_dmul, a.st_level2, a.st_dmul2);
    cpx<R> w[E];
    bool ok[E];
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int64_t k = (int64_t)tl + (int64_t)T * m;
      const int64_t idx = k * a.st_kmul + base;
      ok[m] = have && idx < a.st_limit;
      w[m] = post[ok[m] ? idx : 0];
    }
    mids<R, N, SIGN, 1, E>(line, a, twb, tl);
    __syncthreads();
    cpx<R> u[E];
    stage_last_regs<R, N, SIGN, E>(a, twb, line, tl, u);
    cpx<R>* dst = reinterpret_cast<cpx<R>*>(a.out) + l.out;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int64_t k = (int64_t)tl + (int64_t)T * m;
      if (ok[m]) st(dst + k * a.oks, cmul(u[m], w[m]), nt);
    }
  }
}

// Pass B.  k-fast on both sides (rows are contiguous); the row's data and
// filter are loaded together, the forward DFT's outputs stay in the
// registers of the thread that multiplies them by the filter and starts the
// inverse DFT from them (element tl + T*m in slot m on both sides).
// E: 16 elements per thread, or 8 (twice the threads per row, radix-8/4
// plans) for rows of 1024..4096 points.
template <typename R, int N, bool TWL, int E>
__global__ __launch_bounds__(512) void blue_rows(const sa_pass_t a) {
  constexpr int T = N / E;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  cpx<R>* lds = reinterpret_cast<cpx<R>*>(lds_raw);
  const int logC = a.logC;
  uint32_t c, tl;
  ident<T>(threadIdx.x, true, logC, c, tl);
  cpx<R>* line = lds + (size_t)c * a.ldsstride;
  const uint32_t b = (blockIdx.x << logC) + c;
  const bool have = b < (uint32_t)a.batch;
  const bool nt = a.nt != 0;

  // Table lookups of the store's conjugate four-step twiddle, issued with
  // the row's loads.
  const PtwRaw<R> q = ptw_fetch<R, N, E / StageInfo<N, Plan<N, E>::n - 1, E>::R, E>(a, b, tl);
  cpx<R> v[E], h[E];
  {
    const Line l = line_of(a, have ? b : 0u);
    const cpx<R>* src = reinterpret_cast<const cpx<R>*>(a.in) + l.in + (int64_t)tl * a.iks;
    const cpx<R>* hf = reinterpret_cast<const cpx<R>*>(a.blu_h) +
                       hook_base(a, have ? b : 0u, a.blu_level, a.blu_hstride, -1, 0) + tl;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      v[m] = have ? ld(src + (int64_t)T * m * a.iks, nt) : cpx<R>{R(0), R(0)};
      h[m] = have ? hf[T * m] : cpx<R>{R(0), R(0)};
    }
  }
  const cpx<R>* twb = twiddle_base<R, TWL>(a, lds);
  stage_first<R, N, -1, E>(v, line, tl);
  mids<R, N, -1, 1, E>(line, a, twb, tl);
  __syncthreads();
  stage_last_regs<R, N, -1, E>(a, twb, line, tl, v);
#pragma unroll
  for (int m = 0; m < E; ++m) v[m] = cmul(v[m], h[m]);
  __syncthreads();  // every forward read of the row precedes the inverse's writes
  stage_first<R, N, +1, E>(v, line, tl);
  mids<R, N, +1, 1, E>(line, a, twb, tl);
  __syncthreads();
  // conj(W_M^{n2 k1}) on the store: the backward stages' put multiplies by
  // the conjugate of the table value.
  stage_last<R, N, +1, true, E>(a, twb, line, tl, store_ctx_from<R, N, E>(a, b, tl, q));
}

template <typename K>
static void set_lds_attr(K* fn) {
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(fn), hipFuncAttributeMaxDynamicSharedMemorySize,
                            160 * 1024);
  (void)hipGetLastError();
}

template <typename R, int N, int HK, bool TWL, int E>
static hipError_t launch_cols_t(const sa_pass_t* p, hipStream_t s) {
  static const bool attr = [] {  // once per instantiation, thread-safe
    set_lds_attr(&blue_cols<R, N, HK, TWL, E>);
    return true;
  }();
  (void)attr;
  const uint32_t nthr = (uint32_t)((N / E) << p->logC);
  const uint64_t tiles = ((uint64_t)p->batch + (1u << p->logC) - 1) >> p->logC;
  const size_t lds = (((size_t)p->ldsstride << p->logC) + (TWL ? (size_t)p->tw_lds : 0)) * sizeof(cpx<R>);
  hipLaunchKernelGGL((blue_cols<R, N, HK, TWL, E>), dim3((uint32_t)tiles), dim3(nthr), lds, s, *p);
  return hipGetLastError();
}

template <typename R, int N, bool TWL, int E>
static hipError_t launch_rows_t(const sa_pass_t* p, hipStream_t s) {
  static const bool attr = [] {
    set_lds_attr(&blue_rows<R, N, TWL, E>);
    return true;
  }();
  (void)attr;
  const uint32_t nthr = (uint32_t)((N / E) << p->logC);
  const uint64_t tiles = ((uint64_t)p->batch + (1u << p->logC) - 1) >> p->logC;
  const size_t lds = (((size_t)p->ldsstride << p->logC) + (TWL ? (size_t)p->tw_lds : 0)) * sizeof(cpx<R>);
  hipLaunchKernelGGL((blue_rows<R, N, TWL, E>), dim3((uint32_t)tiles), dim3(nthr), lds, s, *p);
  return hipGetLastError();
}

template <typename R, int N, int E>
static hipError_t launch_cols_e(const sa_pass_t* p, int hook, hipStream_t s) {
  const bool twl = p->tw_lds > 0;
  if (hook == HK_CHIRP_IN)
    return twl ? launch_cols_t<R, N, HK_CHIRP_IN, true, E>(p, s) : launch_cols_t<R, N, HK_CHIRP_IN, false, E>(p, s);
  return twl ? launch_cols_t<R, N, HK_CHIRP_OUT, true, E>(p, s) : launch_cols_t<R, N, HK_CHIRP_OUT, false, E>(p, s);
}

// 16 elements per thread; 32 only for 32-column tiles of 512-point complex
// float columns (512 threads).
template <typename R, int N>
static hipError_t launch_cols(const sa_pass_t* p, int hook, hipStream_t s) {
  if constexpr (sizeof(R) == 4 && N == 512)
    if (p->emax == 32) return launch_cols_e<R, N, 32>(p, hook, s);
  return p->emax == 16 ? launch_cols_e<R, N, 16>(p, hook, s) : hipErrorInvalidValue;
}

template <typename R, int N>
static hipError_t launch_rows(const sa_pass_t* p, hipStream_t s) {
  if constexpr (N <= 4096)
    if (p->emax == 8)
      return p->tw_lds > 0 ? launch_rows_t<R, N, true, 8>(p, s) : launch_rows_t<R, N, false, 8>(p, s);
  if (p->emax != 16) return hipErrorInvalidValue;
  return p->tw_lds > 0 ? launch_rows_t<R, N, true, 16>(p, s) : launch_rows_t<R, N, false, 16>(p, s);
}

template <typename R>
static hipError_t launch_prec(const sa_pass_t* p, int kind, hipStream_t s) {
  if (kind == SA_BLUE_ROWS) {
    switch (p->n) {
      case 1024: return launch_rows<R, 1024>(p, s);
      case 2048: return launch_rows<R, 2048>(p, s);
      case 4096: return launch_rows<R, 4096>(p, s);
      case 8192: return launch_rows<R, 8192>(p, s);
      default: return hipErrorInvalidValue;
    }
  }
  const int hook = kind == SA_BLUE_COLS_IN ? HK_CHIRP_IN : HK_CHIRP_OUT;
  switch (p->n) {
    case 32: return launch_cols<R, 32>(p, hook, s);
    case 64: return launch_cols<R, 64>(p, hook, s);
    case 128: return launch_cols<R, 128>(p, hook, s);
    case 256: return launch_cols<R, 256>(p, hook, s);
    case 512: return launch_cols<R, 512>(p, hook, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace fx
}  // namespace sa

extern "C" int sa_blue_elems(int n, int prec, int kind, int lines) {
  // Complex-float 512-point columns in 32-column tiles hold 32 elements per
  // thread (512 threads); every other tile 16.
  return (kind != SA_BLUE_ROWS && prec == 0 && n == 512 && lines > 16) ? 32 : 16;
}

extern "C" int sa_blue_supported(int n, int kind) {
  if (kind == SA_BLUE_ROWS) return n == 1024 || n == 2048 || n == 4096 || n == 8192;
  if (kind == SA_BLUE_COLS_IN || kind == SA_BLUE_COLS_OUT)
    return n == 32 || n == 64 || n == 128 || n == 256 || n == 512;
  return 0;
}

extern "C" int sa_launch_blue(const sa_pass_t* p, int prec, int kind, void* stream) {
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (!sa_blue_supported(p->n, kind)) return (int)hipErrorInvalidValue;
  return prec == 0 ? (int)sa::fx::launch_prec<float>(p, kind, s) : (int)sa::fx::launch_prec<double>(p, kind, s);
}
