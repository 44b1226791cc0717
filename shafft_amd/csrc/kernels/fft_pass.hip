// Generic LDS-resident FFT pass for gfx950.
//
// One workgroup owns C lines of one pass (see pass_desc.h for the
// addressing).  The lines are pulled from HBM into LDS with a lane mapping
// chosen per side so that every wavefront load/store touches contiguous
// bytes: "k-fast" when the FFT axis itself is unit-stride, "line-fast" when
// neighbouring lines are adjacent in memory (strided axes, fused transposes).
// The transform is a mixed-radix Stockham autosort (radices 2,3,4,5,7,8,16)
// run in place in LDS; each stage holds a thread's butterflies in registers
// between two barriers.  Optional fused element-wise work on the way out:
// four-step twiddle, normalisation scale, Bluestein chirps.
//
// Every pass reads and writes each element of its lines exactly once, so the
// pass is HBM-bound: algorithmic bytes per pass = 2 * batch * n * sizeof(T).
#include "fft_device.hpp"
#include "pass_desc.h"

namespace sa {

// Register budget per thread per stage, in complex elements.
constexpr int EMAX = 16;

template <typename R, int RAD, int SIGN>
__device__ __forceinline__ void stockham_stage(cpx<R>* line, int n, int Ns, sa_fastdiv_t nsd,
                                               const cpx<R>* __restrict__ tw, int tl, int T) {
  constexpr int MAXB = (EMAX + RAD - 1) / RAD;
  const int nb = n / RAD;
  const FastDiv fd{nsd.d, nsd.m, nsd.s};
  cpx<R> v[MAXB][RAD];
#pragma unroll
  for (int i = 0; i < MAXB; ++i) {
    const int jb = tl + i * T;
    if (jb < nb) {
#pragma unroll
      for (int r = 0; r < RAD; ++r) v[i][r] = line[jb + r * nb];
      if (Ns > 1) {
        const int q = (int)fdiv((uint32_t)jb, fd);
        const int rem = jb - q * Ns;
        const cpx<R>* w = tw + rem * (RAD - 1);
#pragma unroll
        for (int r = 1; r < RAD; ++r) v[i][r] = cmul_dir<SIGN>(v[i][r], w[r - 1]);
      }
      dft<RAD, SIGN>(v[i]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < MAXB; ++i) {
    const int jb = tl + i * T;
    if (jb < nb) {
      const int q = (int)fdiv((uint32_t)jb, fd);
      const int rem = jb - q * Ns;
      const int idx = q * Ns * RAD + rem;
#pragma unroll
      for (int r = 0; r < RAD; ++r) line[idx + r * Ns] = v[i][r];
    }
  }
  __syncthreads();
}

template <typename R, int SIGN>
__device__ __forceinline__ void run_stages(const sa_pass_t& a, cpx<R>* line, int tl) {
  const cpx<R>* tw = reinterpret_cast<const cpx<R>*>(a.tw);
  for (int s = 0; s < a.nstages; ++s) {
    const cpx<R>* tws = tw + a.twoff[s];
    switch (a.radix[s]) {
      case 2: stockham_stage<R, 2, SIGN>(line, a.n, a.ns[s], a.nsdiv[s], tws, tl, a.T); break;
      case 3: stockham_stage<R, 3, SIGN>(line, a.n, a.ns[s], a.nsdiv[s], tws, tl, a.T); break;
      case 4: stockham_stage<R, 4, SIGN>(line, a.n, a.ns[s], a.nsdiv[s], tws, tl, a.T); break;
      case 5: stockham_stage<R, 5, SIGN>(line, a.n, a.ns[s], a.nsdiv[s], tws, tl, a.T); break;
      case 7: stockham_stage<R, 7, SIGN>(line, a.n, a.ns[s], a.nsdiv[s], tws, tl, a.T); break;
      case 8: stockham_stage<R, 8, SIGN>(line, a.n, a.ns[s], a.nsdiv[s], tws, tl, a.T); break;
      case 16: stockham_stage<R, 16, SIGN>(line, a.n, a.ns[s], a.nsdiv[s], tws, tl, a.T); break;
      default: break;
    }
  }
}

__device__ __forceinline__ uint32_t fdiv_d(uint32_t x, sa_fastdiv_t f) {
  return fdiv(x, FastDiv{f.d, f.m, f.s});
}

// Offsets of line b (global line index) for the in- and output side.
struct LineOff {
  int64_t in, out;
  uint32_t d[3];
};

__device__ __forceinline__ LineOff line_offsets(const sa_pass_t& a, uint32_t b) {
  LineOff o;
  const uint32_t q2 = fdiv_d(b, a.n2div);
  const uint32_t b2 = b - q2 * (uint32_t)a.n2;
  const uint32_t b0 = fdiv_d(q2, a.n1div);
  const uint32_t b1 = q2 - b0 * (uint32_t)a.n1;
  o.in = (int64_t)b0 * a.is0 + (int64_t)b1 * a.is1 + (int64_t)b2 * a.is2;
  o.out = (int64_t)b0 * a.os0 + (int64_t)b1 * a.os1 + (int64_t)b2 * a.os2;
  o.d[0] = b0;
  o.d[1] = b1;
  o.d[2] = b2;
  return o;
}

template <typename R, int SIGN, bool BLUE>
__global__ __launch_bounds__(512) void fft_lds_pass(const sa_pass_t a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  cpx<R>* lds = reinterpret_cast<cpx<R>*>(lds_raw);

  const uint32_t t = threadIdx.x;
  const uint32_t nthr = blockDim.x;
  const uint32_t line0 = blockIdx.x * (uint32_t)a.C;
  const int n = a.n;

  // ---- load: HBM -> LDS ----------------------------------------------------
  {
    uint32_t c, k0;
    if (a.ld_kfast) {
      c = fdiv_d(t, a.Tdiv);
      k0 = t - c * (uint32_t)a.T;
    } else {
      k0 = fdiv_d(t, a.Cdiv);
      c = t - k0 * (uint32_t)a.C;
    }
    const uint32_t kstep = nthr / (uint32_t)a.C;
    const uint32_t b = line0 + c;
    cpx<R>* line = lds + (size_t)c * a.ldsstride;
    if (b < (uint64_t)a.batch) {
      const LineOff o = line_offsets(a, b);
      const cpx<R>* __restrict__ src = reinterpret_cast<const cpx<R>*>(a.in) + o.in;
      const bool plain = a.pre == nullptr && a.ld_level < 0 && a.ld_kmul == 1 && a.ld_limit >= n;
      cpx<R> buf[EMAX];
      if (plain) {
#pragma unroll
        for (int i = 0; i < EMAX; ++i) {
          const uint32_t k = k0 + i * kstep;
          if (k < (uint32_t)n) buf[i] = src[(int64_t)k * a.iks];
        }
      } else {
        const cpx<R>* __restrict__ pre = reinterpret_cast<const cpx<R>*>(a.pre);
        const int64_t base = (a.ld_level >= 0 ? (int64_t)o.d[a.ld_level] * a.ld_dmul : 0) +
                             (a.ld_level2 >= 0 ? (int64_t)o.d[a.ld_level2] * a.ld_dmul2 : 0);
#pragma unroll
        for (int i = 0; i < EMAX; ++i) {
          const uint32_t k = k0 + i * kstep;
          if (k < (uint32_t)n) {
            const int64_t m = (int64_t)k * a.ld_kmul + base;
            cpx<R> v = {R(0), R(0)};
            if (m < a.ld_limit) {
              v = src[(int64_t)k * a.iks];
              if (pre) v = cmul(v, pre[m]);
            }
            buf[i] = v;
          }
        }
      }
#pragma unroll
      for (int i = 0; i < EMAX; ++i) {
        const uint32_t k = k0 + i * kstep;
        if (k < (uint32_t)n) line[k] = buf[i];
      }
    }
  }
  __syncthreads();

  // ---- Stockham stages in LDS ------------------------------------------------
  const uint32_t cs = fdiv_d(t, a.Tdiv);
  const int tl = (int)(t - cs * (uint32_t)a.T);
  cpx<R>* sline = lds + (size_t)cs * a.ldsstride;
  if constexpr (!BLUE) {
    run_stages<R, SIGN>(a, sline, tl);
  } else {
    run_stages<R, -1>(a, sline, tl);
    const cpx<R>* __restrict__ h = reinterpret_cast<const cpx<R>*>(a.blu_h);
    for (int k = tl; k < n; k += a.T) sline[k] = cmul(sline[k], h[k]);
    __syncthreads();
    run_stages<R, +1>(a, sline, tl);
  }

  // ---- store: LDS -> HBM ---------------------------------------------------
  {
    uint32_t c, k0;
    if (a.st_kfast) {
      c = fdiv_d(t, a.Tdiv);
      k0 = t - c * (uint32_t)a.T;
    } else {
      k0 = fdiv_d(t, a.Cdiv);
      c = t - k0 * (uint32_t)a.C;
    }
    const uint32_t kstep = nthr / (uint32_t)a.C;
    const uint32_t b = line0 + c;
    if (b < (uint64_t)a.batch) {
      const cpx<R>* line = lds + (size_t)c * a.ldsstride;
      const LineOff o = line_offsets(a, b);
      cpx<R>* __restrict__ dst = reinterpret_cast<cpx<R>*>(a.out) + o.out;
      const cpx<R>* __restrict__ post = reinterpret_cast<const cpx<R>*>(a.post);
      const bool do_scale = a.scale != 1.0;
      const R sc = (R)a.scale;
      if (a.ptw_level >= 0) {
        // W_N^{(digit * k) mod N}, walked incrementally along k.
        const uint64_t N = (uint64_t)a.ptw_N;
        const uint64_t dig = o.d[a.ptw_level];
        uint64_t e = (dig * k0) % N;
        const uint64_t step = (dig * kstep) % N;
        const cpx<R>* __restrict__ lo = reinterpret_cast<const cpx<R>*>(a.ptw_lo);
        const cpx<R>* __restrict__ hi = reinterpret_cast<const cpx<R>*>(a.ptw_hi);
        const uint64_t mask = (1ull << a.ptw_shift) - 1;
#pragma unroll
        for (int i = 0; i < EMAX; ++i) {
          const uint32_t k = k0 + i * kstep;
          if (k < (uint32_t)n) {
            const cpx<R> w = cmul(hi[e >> a.ptw_shift], lo[e & mask]);
            cpx<R> v = cmul_dir<SIGN>(line[k], w);
            if (do_scale) v = cscale(v, sc);
            dst[(int64_t)k * a.oks] = v;
          }
          e += step;
          if (e >= N) e -= N;
        }
      } else if (post == nullptr && a.st_level < 0 && a.st_kmul == 1 && a.st_limit >= n) {
#pragma unroll
        for (int i = 0; i < EMAX; ++i) {
          const uint32_t k = k0 + i * kstep;
          if (k < (uint32_t)n) {
            cpx<R> v = line[k];
            if (do_scale) v = cscale(v, sc);
            dst[(int64_t)k * a.oks] = v;
          }
        }
      } else {
        const int64_t base = (a.st_level >= 0 ? (int64_t)o.d[a.st_level] * a.st_dmul : 0) +
                             (a.st_level2 >= 0 ? (int64_t)o.d[a.st_level2] * a.st_dmul2 : 0);
#pragma unroll
        for (int i = 0; i < EMAX; ++i) {
          const uint32_t k = k0 + i * kstep;
          if (k < (uint32_t)n) {
            const int64_t m = (int64_t)k * a.st_kmul + base;
            if (m < a.st_limit) {
              cpx<R> v = line[k];
              if (post) v = cmul(v, post[m]);
              if (do_scale) v = cscale(v, sc);
              dst[(int64_t)k * a.oks] = v;
            }
          }
        }
      }
    }
  }
}

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

template <typename R, int SIGN, bool BLUE>
static hipError_t launch_t(const sa_pass_t* p, hipStream_t st) {
  const uint32_t nthr = (uint32_t)(p->C * p->T);
  const uint64_t nblk = ((uint64_t)p->batch + p->C - 1) / p->C;
  const size_t lds = (size_t)p->C * p->ldsstride * sizeof(cpx<R>);
  static bool attr_done = false;
  if (!attr_done) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&fft_lds_pass<R, SIGN, BLUE>),
                        hipFuncAttributeMaxDynamicSharedMemorySize, sa_max_lds_bytes());
    attr_done = true;
  }
  hipLaunchKernelGGL((fft_lds_pass<R, SIGN, BLUE>), dim3((uint32_t)nblk), dim3(nthr), lds, st, *p);
  return hipGetLastError();
}

}  // namespace sa

extern "C" int sa_max_lds_bytes(void) { return 160 * 1024; }

extern "C" int sa_launch_pass(const sa_pass_t* p, int prec, int sign, void* stream) {
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const bool blue = p->blu_h != nullptr;
  if (prec == 0) {
    if (blue) return (int)sa::launch_t<float, -1, true>(p, st);
    return sign < 0 ? (int)sa::launch_t<float, -1, false>(p, st)
                    : (int)sa::launch_t<float, +1, false>(p, st);
  }
  if (blue) return (int)sa::launch_t<double, -1, true>(p, st);
  return sign < 0 ? (int)sa::launch_t<double, -1, false>(p, st)
                  : (int)sa::launch_t<double, +1, false>(p, st);
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
