// Generic LDS-resident FFT pass for gfx950.
//
// One workgroup owns C lines of one pass (see pass_desc.h for the
// addressing).  Lines move between HBM and the workgroup with a lane mapping
// chosen per side so that every wavefront load/store touches contiguous
// bytes: "k-fast" when the FFT axis itself is unit-stride, "line-fast" when
// neighbouring lines are adjacent in memory (strided axes, fused transposes).
// The transform is a mixed-radix Stockham autosort (radices 2,3,4,5,7,8,16).
// When the line length is 16 x (threads per line), the first stage runs on
// the registers the load filled and the last stage leaves its outputs in the
// registers the store drains, so LDS only carries the inner exchanges.  LDS
// rows are padded by one element every 16 (index k -> k + k/16) so the
// stride-16 butterfly scatter of the first stage is bank-conflict free.
// Optional fused element-wise work on the way in/out: four-step twiddle,
// scale, Bluestein chirps and filter.
//
// Every pass reads and writes each element of its lines once, so the pass is
// HBM-bound: algorithmic bytes per pass = 2 * batch * n * sizeof(T).
#include "pass_tile.hpp"

namespace sa {

// Tile of this workgroup.  Workgroups are dealt to the 8 XCDs round robin
// (workgroup b on XCD b % 8).  With xcd_order set, XCD x takes a contiguous
// range of tiles instead (workgroup x + 8 i -> tile start_x + i, a bijection
// on [0, grid)), so tiles next to each other run at the same time on one
// XCD and the partial cache lines at the edges of their row segments meet
// in one L2 instead of being fetched and written back from two.
__device__ __forceinline__ uint32_t tile_of_block(const sa_pass_t& a) {
  const uint32_t b = blockIdx.x;
  if (!a.xcd_order) return b;
  const uint32_t q = gridDim.x >> 3, r = gridDim.x & 7u, x = b & 7u;
  return x * q + (x < r ? x : r) + (b >> 3);
}

// Register budget and occupancy.  With the stage loop's LDS addresses
// computed inside each stage (stage_lane), the LDS-mode kernels need ~94
// (complex float, radices up to 21) / ~151 (complex double, up to 28) VGPRs,
// 5 / 3 waves per SIMD.  The planner's tiles hold at most 64 KiB of data
// (512 threads complex float, 256 complex double), so several independent
// tiles share a CU and one's HBM loads overlap another's LDS stages (the
// stages issue ~100-160 VALU instructions per element over 3 stages, load
// and store, by SQ counters; how far the two workgroups' memory and LDS
// phases overlap sets the pass time more than the instruction count).
// E: elements per thread (8 only for the in-LDS Bluestein rows of the
// three-pass long Bluestein: twice the threads on a pass of few, long rows).
// PAD: LDS rows padded (k -> k + k/16); unpadded for passes whose accesses
// spread over the banks without it (the planner's choice, a.lds_pad).
template <int SIGN, int MODE, bool TWL, int E = EMAX, bool PAD = true>
__global__ __launch_bounds__(512) void fft_lds_pass_f(const sa_pass_t a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  pass_tile<float, SIGN, MODE, 0, TWL, E, PAD>(a, tile_of_block(a), reinterpret_cast<cpx<float>*>(lds_raw));
}

template <int SIGN, int MODE, bool TWL, int E = EMAX, bool PAD = true>
__global__ __launch_bounds__(512) void fft_lds_pass_d(const sa_pass_t a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  pass_tile<double, SIGN, MODE, 0, TWL, E, PAD>(a, tile_of_block(a), reinterpret_cast<cpx<double>*>(lds_raw));
}

template <typename R, int SIGN, int MODE, bool TWL, int E = EMAX, bool PAD = true>
struct PassKernel;
template <int SIGN, int MODE, bool TWL, int E, bool PAD>
struct PassKernel<float, SIGN, MODE, TWL, E, PAD> {
  static const void* fn() { return reinterpret_cast<const void*>(&fft_lds_pass_f<SIGN, MODE, TWL, E, PAD>); }
};
template <int SIGN, int MODE, bool TWL, int E, bool PAD>
struct PassKernel<double, SIGN, MODE, TWL, E, PAD> {
  static const void* fn() { return reinterpret_cast<const void*>(&fft_lds_pass_d<SIGN, MODE, TWL, E, PAD>); }
};

template <typename R>
__global__ __launch_bounds__(256) void scale_kernel(cpx<R>* __restrict__ x, R f, uint64_t count) {
  const uint64_t stride = (uint64_t)blockDim.x * gridDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
    cpx<R> v = x[i];
    v.x *= f;
    v.y *= f;
    x[i] = v;
  }
}

// TWL: the stage twiddles are staged in LDS behind the tile (a.tw_lds > 0).
template <typename R, int SIGN, int MODE, bool TWL, int E = EMAX, bool PAD = true>
static hipError_t launch_t(const sa_pass_t* p, hipStream_t st) {
  const uint32_t nthr = (uint32_t)(p->C * p->T);
  const uint64_t nblk = ((uint64_t)p->batch + p->C - 1) / p->C;
  const size_t lds = ((size_t)p->C * p->ldsstride + (TWL ? (size_t)p->tw_lds : 0)) * sizeof(cpx<R>);
  static const bool attr = [] {  // once per instantiation, thread-safe
    (void)hipFuncSetAttribute(PassKernel<R, SIGN, MODE, TWL, E, PAD>::fn(),
                              hipFuncAttributeMaxDynamicSharedMemorySize, sa_max_lds_bytes());
    (void)hipGetLastError();
    return true;
  }();
  (void)attr;
  if constexpr (sizeof(R) == 4)
    hipLaunchKernelGGL((fft_lds_pass_f<SIGN, MODE, TWL, E, PAD>), dim3((uint32_t)nblk), dim3(nthr), lds, st, *p);
  else
    hipLaunchKernelGGL((fft_lds_pass_d<SIGN, MODE, TWL, E, PAD>), dim3((uint32_t)nblk), dim3(nthr), lds, st, *p);
  return hipGetLastError();
}

template <typename R, int SIGN, int MODE>
static hipError_t launch_tw(const sa_pass_t* p, hipStream_t st) {
  if (!p->lds_pad)
    return p->tw_lds > 0 ? launch_t<R, SIGN, MODE, true, EMAX, false>(p, st)
                         : launch_t<R, SIGN, MODE, false, EMAX, false>(p, st);
  return p->tw_lds > 0 ? launch_t<R, SIGN, MODE, true>(p, st) : launch_t<R, SIGN, MODE, false>(p, st);
}

template <typename R>
static hipError_t launch_prec(const sa_pass_t* p, int sign, hipStream_t st) {
  if (p->blu_h) {
    const bool twl = p->tw_lds > 0;
    if constexpr (sizeof(R) == 4)  // 8 elements per thread: complex float only
      if (p->emax == 8)
        return twl ? launch_t<R, -1, MODE_BLUE, true, 8>(p, st) : launch_t<R, -1, MODE_BLUE, false, 8>(p, st);
    return twl ? launch_t<R, -1, MODE_BLUE, true>(p, st) : launch_t<R, -1, MODE_BLUE, false>(p, st);
  }
  const bool regs = p->reg_first && p->reg_last;
  if (sign < 0) return regs ? launch_tw<R, -1, MODE_REGS>(p, st) : launch_tw<R, -1, MODE_LDS>(p, st);
  return regs ? launch_tw<R, +1, MODE_REGS>(p, st) : launch_tw<R, +1, MODE_LDS>(p, st);
}

}  // namespace sa

extern "C" int sa_max_lds_bytes(void) { return 160 * 1024; }

namespace sa {
template <typename R>
static const void* pass_kernel(const sa_pass_t* p, bool twl) {
  if (p->blu_h) {
    if constexpr (sizeof(R) == 4)
      if (p->emax == 8)
        return twl ? PassKernel<R, -1, MODE_BLUE, true, 8>::fn() : PassKernel<R, -1, MODE_BLUE, false, 8>::fn();
    return twl ? PassKernel<R, -1, MODE_BLUE, true>::fn() : PassKernel<R, -1, MODE_BLUE, false>::fn();
  }
  if (!p->lds_pad) {
    if (p->reg_first && p->reg_last)
      return twl ? PassKernel<R, -1, MODE_REGS, true, EMAX, false>::fn()
                 : PassKernel<R, -1, MODE_REGS, false, EMAX, false>::fn();
    return twl ? PassKernel<R, -1, MODE_LDS, true, EMAX, false>::fn()
               : PassKernel<R, -1, MODE_LDS, false, EMAX, false>::fn();
  }
  if (p->reg_first && p->reg_last)
    return twl ? PassKernel<R, -1, MODE_REGS, true>::fn() : PassKernel<R, -1, MODE_REGS, false>::fn();
  return twl ? PassKernel<R, -1, MODE_LDS, true>::fn() : PassKernel<R, -1, MODE_LDS, false>::fn();
}
}  // namespace sa

extern "C" int sa_pass_occupancy(const sa_pass_t* p, int prec, int twl) {
  const size_t esz = prec ? 16 : 8;
  const size_t lds = ((size_t)p->C * p->ldsstride + (twl ? (size_t)p->tw_lds : 0)) * esz;
  const void* f = prec ? sa::pass_kernel<double>(p, twl != 0) : sa::pass_kernel<float>(p, twl != 0);
  (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, sa_max_lds_bytes());
  (void)hipGetLastError();
  int blocks = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, f, p->C * p->T, lds) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return blocks;
}

extern "C" int sa_launch_pass(const sa_pass_t* p, int prec, int sign, void* stream) {
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  return prec == 0 ? (int)sa::launch_prec<float>(p, sign, st) : (int)sa::launch_prec<double>(p, sign, st);
}

extern "C" int sa_launch_scale(void* data, double f, uint64_t count, int prec, void* stream) {
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (count == 0) return 0;
  uint64_t blocks = (count + 255) / 256;
  if (blocks > 65535ull * 8) blocks = 65535ull * 8;
  if (prec == 0)
    hipLaunchKernelGGL(sa::scale_kernel<float>, dim3((uint32_t)blocks), dim3(256), 0, st,
                       reinterpret_cast<sa::cpx<float>*>(data), (float)f, count);
  else
    hipLaunchKernelGGL(sa::scale_kernel<double>, dim3((uint32_t)blocks), dim3(256), 0, st,
                       reinterpret_cast<sa::cpx<double>*>(data), f, count);
  return (int)hipGetLastError();
}
