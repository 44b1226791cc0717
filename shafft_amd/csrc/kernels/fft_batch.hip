// Persistent batched FFT kernel for gfx950: many transforms of uneven shape
// in one launch (dependency mode) or one launch per pass level (level mode,
// all launches in one captured graph).
//
// The host flattens every pass of every transform into a work list of
// (pass, tile) items in level-major order (all first passes, then all second
// passes, ...).  Workgroups pull items from a device-side queue head with
// one atomic per item.  An item whose pass depends on an earlier pass of
// the same transform waits (one lane spins on an agent-scope acquire load)
// until that pass's completed-tile counter is full; a finished tile makes
// its stores visible with an agent-scope release before counting itself.
// Every dependency of an item was handed out before the item itself, to a
// workgroup that is already running, so the queue always drains.  The
// agent-scope release is an L2 write-back on gfx950, so level mode, where
// the kernel boundary orders the levels, is usually faster; the autotuner
// times both.
#include "fft_device.hpp"
#include "pass_desc.h"
#include "pass_tile.hpp"

namespace sa {

// Clears the per-pass completion counters and the work-queue head.  The
// wait-timeout flag is not cleared here: it stays set across launches (and
// graph replays) until the host has read it.
__global__ void batch_clear(uint32_t* done, uint32_t n, uint32_t* head, uint32_t queues) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) done[i] = 0u;
  if (i < queues) head[i] = 0u;
}

template <typename R, int MX>
__global__ __launch_bounds__(512) void fft_batch_kernel(const sa_batch_t b) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  __shared__ uint32_t s_item;
  cpx<R>* lds = reinterpret_cast<cpx<R>*>(lds_raw);
  // Every branch and loop bound below that encloses a barrier is made
  // wave-uniform with readfirstlane (see uni() in pass_tile.hpp).
  // Queue q holds items [q * per, min((q + 1) * per, nitems)); a workgroup
  // starts on its XCD's queue and moves on to the next when that is used up,
  // so the loop ends after at most nitems + queues taken counters.
  const uint32_t nq = b.queues > 1 ? b.queues : 1u;
  const uint32_t per = (b.nitems + nq - 1) / nq;
  uint32_t q = nq > 1 ? (blockIdx.x % nq) : 0u, tried = 0;
  for (;;) {
    if (threadIdx.x == 0)
      s_item = __hip_atomic_fetch_add(b.head + q, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const uint32_t k = (uint32_t)uni((int)s_item);
    __syncthreads();
    const uint32_t lo = q * per, hi = min(lo + per, b.nitems);
    if (lo + k >= hi) {  // this queue is used up
      if (++tried >= nq) break;
      q = q + 1 == nq ? 0u : q + 1;
      continue;
    }
    const uint32_t it = lo + k;
    const uint32_t pass = (uint32_t)uni((int)b.items[it].pass);
    const uint32_t tile = (uint32_t)uni((int)b.items[it].tile);
    const int dep = b.independent ? -1 : uni(b.pass_dep[pass]);
    if (dep >= 0) {
      if (threadIdx.x == 0) {
        // Bounded wait (wall clock): a wait that outlives wait_ticks flags
        // the launch as failed instead of holding the GPU.
        const uint32_t need = b.pass_tiles[dep];
        const uint64_t t0 = wall_clock64();
        while (__hip_atomic_load(&b.done[dep], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < need) {
          __builtin_amdgcn_s_sleep(2);
          if (wall_clock64() - t0 > b.wait_ticks) {
            __hip_atomic_store(b.err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
          }
        }
      }
      __syncthreads();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    const sa_pass_t& a = b.passes[pass];
    const int kind = uni(a.blu_h != nullptr ? 2 : (b.pass_sign[pass] < 0 ? 0 : 1));
    if (kind == 2)
      pass_tile<R, -1, MODE_BLUE>(a, tile, lds);
    else if (kind == 0)
      pass_tile<R, -1, MODE_LDS, MX>(a, tile, lds);
    else
      pass_tile<R, +1, MODE_LDS, MX>(a, tile, lds);
    if (!b.independent) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      __syncthreads();
      if (threadIdx.x == 0)
        __hip_atomic_fetch_add(&b.done[pass], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
  }
}

}  // namespace sa

extern "C" int sa_launch_batch(const sa_batch_t* b, int prec, int grid, int block, size_t lds,
                               void* stream) {
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const uint32_t n = b->npasses;
  const uint32_t nq = b->queues > 1 ? b->queues : 1u;
  if (nq != 1 && (nq != 8 || !b->independent)) return (int)hipErrorInvalidValue;  // see sa_batch_t
  if (!b->precleared) {
    const uint32_t m = n > nq ? n : nq;
    hipLaunchKernelGGL(sa::batch_clear, dim3((m + 255) / 256), dim3(256), 0, st, b->done, n, b->head, nq);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return (int)e;
  }
  // 160 KiB minus room for the kernel's static LDS (the queue item slot);
  // set once per variant (thread-safe static init).
  static const bool attr = [] {
    const void* fs[4] = {reinterpret_cast<const void*>(&sa::fft_batch_kernel<float, 0>),
                         reinterpret_cast<const void*>(&sa::fft_batch_kernel<float, 1>),
                         reinterpret_cast<const void*>(&sa::fft_batch_kernel<float, 2>),
                         reinterpret_cast<const void*>(&sa::fft_batch_kernel<double, 0>)};
    for (const void* f : fs)
      (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 159 * 1024);
    (void)hipGetLastError();
    return true;
  }();
  (void)attr;
  if (prec && b->mfma) return (int)hipErrorInvalidValue;  // split stages are complex-float only
  if ((block & 63) != 0) return (int)hipErrorInvalidValue;  // matrix-core stages need whole waves
  if (b->mfma) {  // the stages' constant operands, ahead of the kernel on its stream
    hipLaunchKernelGGL(sa::mx_table_init, dim3(1), dim3(64), 0, st);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return (int)e;
  }
  const dim3 g((uint32_t)grid), t((uint32_t)block);
  if (prec)
    hipLaunchKernelGGL((sa::fft_batch_kernel<double, 0>), g, t, lds, st, *b);
  else if (b->mfma == 1)
    hipLaunchKernelGGL((sa::fft_batch_kernel<float, 1>), g, t, lds, st, *b);
  else if (b->mfma == 2)
    hipLaunchKernelGGL((sa::fft_batch_kernel<float, 2>), g, t, lds, st, *b);
  else
    hipLaunchKernelGGL((sa::fft_batch_kernel<float, 0>), g, t, lds, st, *b);
  return (int)hipGetLastError();
}

extern "C" int sa_batch_occupancy(int prec, int block, size_t lds) {
  const void* f = prec ? reinterpret_cast<const void*>(&sa::fft_batch_kernel<double, 0>)
                       : reinterpret_cast<const void*>(&sa::fft_batch_kernel<float, 0>);
  (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 159 * 1024);
  (void)hipGetLastError();
  int blocks = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, f, block, lds) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return blocks;
}
