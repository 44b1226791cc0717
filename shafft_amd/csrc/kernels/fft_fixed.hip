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
#include "fixed_stages.hpp"

namespace sa {
namespace fx {

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
  if (a.xcd_order == 1 && (gridDim.x & 7u) == 0) {
    const uint32_t x = blockIdx.x & 7u, per = (tiles + 7u) >> 3;
    tile = x * per + (blockIdx.x >> 3);
    step = gridDim.x >> 3;
    end = min(tiles, (x + 1u) * per);
  }
  // xcd_order >= 2: the tiles form rows of rlen = xcd_order tiles and are
  // walked down the rows first, so the tiles in flight at once spread over
  // many rows (the planner sets it only when rlen divides the tile count;
  // t -> (t % rows) * rlen + t / rows is then a bijection on the tiles).
  const bool cm = a.xcd_order >= 2;
  const uint32_t rlen = cm ? (uint32_t)a.xcd_order : 1u, rows = cm ? tiles / rlen : 1u;
  auto at = [&](uint32_t t) { return cm ? (t % rows) * rlen + t / rows : t; };
  cpx<R> v[E];
  if (tile < end) load_tile<R, N>(a, (at(tile) << logC) + c, tl, v);
  while (tile < end) {
    stage_first<R, N, SIGN>(v, line, tl);
    const uint32_t next = tile + step;
    if (next < end) load_tile<R, N>(a, (at(next) << logC) + c, tl, v);  // prefetch
    mids<R, N, SIGN, 1>(line, a, twb, tl);
    __syncthreads();
    if (same) {
      const StoreCtx<R> ctx =
          store_ctx<R, N, E / StageInfo<N, Plan<N, E>::n - 1, E>::R>(a, (at(tile) << logC) + c, tl);
      stage_last<R, N, SIGN, true>(a, twb, line, tl, ctx);
    } else {
      // Transposing pass: natural order through LDS, then the store mapping.
      const StoreCtx<R> dummy = {};
      stage_last<R, N, SIGN, false>(a, twb, line, tl, dummy);
      __syncthreads();
      const StoreCtx<R> ctx = store_ctx<R, N, 1>(a, (at(tile) << logC) + sc, stl);
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
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, f, nthr, lds) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return blocks;
}
