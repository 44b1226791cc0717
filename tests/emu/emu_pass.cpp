// CPU emulation of the pass launch ABI (kernels/pass_desc.h) and of the few
// HIP runtime calls the host plan code makes.  Test infrastructure only: it
// lets the planner's pass programs (digits, strides, four-step twiddles,
// fused transposes, Bluestein hooks, buffer slots) be checked against numpy
// on a machine without a GPU.  The Stockham stages are replayed with the
// plan's own twiddle tables and radices; butterflies are direct DFTs.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <complex>
#include <cstdlib>
#include <algorithm>
#include <array>
#include <iterator>
#include <cstdio>
#include <functional>
#include <map>
#include <mutex>
#include <set>
#include <vector>
#include <cstring>
#include <algorithm>
#include <vector>

#include "../../shafft_amd/csrc/kernels/pass_desc.h"

namespace {

// Allocation registry so every emulated access can be bounds-checked.
// Plans may be built on other host threads while a pass is replayed, so the
// registry is guarded.
std::map<uintptr_t, size_t>& allocs() {
  static std::map<uintptr_t, size_t> m;
  return m;
}
std::mutex& allocs_mu() {
  static std::mutex mu;
  return mu;
}

[[noreturn]] void oob(const char* what, const void* base, int64_t idx) {
  std::fprintf(stderr, "EMU OUT OF BOUNDS: %s base=%p index=%lld\n", what, base, (long long)idx);
  std::abort();
}

// Checks that element idx (of size esz) of the array at base lies inside one
// registered allocation.
void chk(const char* what, const void* base, int64_t idx, size_t esz) {
  const uintptr_t a = (uintptr_t)base + (uintptr_t)(idx * (int64_t)esz);
  if (idx < 0) oob(what, base, idx);
  std::lock_guard<std::mutex> lk(allocs_mu());
  auto& m = allocs();
  auto it = m.upper_bound(a);
  if (it == m.begin()) oob(what, base, idx);
  --it;
  if (a + esz > it->first + it->second || (uintptr_t)base < it->first) oob(what, base, idx);
}

template <typename R>
using C = std::complex<R>;

template <typename R>
C<R> conj_if(C<R> w, int sign) {
  return sign < 0 ? w : std::conj(w);
}

template <typename R>
void stages(const sa_pass_t& p, std::vector<C<R>>& d, int sign) {
  const int n = p.n;
  const C<R>* tw = reinterpret_cast<const C<R>*>(p.tw);
  std::vector<C<R>> out(n);
  for (int s = 0; s < p.nstages; ++s) {
    const int Rr = p.radix[s], Ns = p.ns[s], nb = n / Rr;
    const C<R>* tws = tw + p.twoff[s];
    for (int jb = 0; jb < nb; ++jb) {
      const int q = jb / Ns, rem = jb % Ns;
      std::vector<C<R>> v(Rr), V(Rr);
      for (int r = 0; r < Rr; ++r) {
        v[r] = d.at(jb + r * nb);
        if (Ns > 1 && r > 0) {
          chk("tw", tws, rem * (Rr - 1) + r - 1, sizeof(C<R>));
          v[r] *= conj_if(tws[rem * (Rr - 1) + r - 1], sign);
        }
      }
      for (int k = 0; k < Rr; ++k) {
        C<long double> acc = 0;
        for (int r = 0; r < Rr; ++r) {
          const long double a = sign * 2.0L * 3.14159265358979323846264338327950288L *
                                (long double)((r * k) % Rr) / Rr;
          acc += C<long double>(v[r].real(), v[r].imag()) * C<long double>(cosl(a), sinl(a));
        }
        V[k] = C<R>((R)acc.real(), (R)acc.imag());
      }
      const int idx = q * Ns * Rr + rem;
      for (int r = 0; r < Rr; ++r) out.at(idx + r * Ns) = V[r];
    }
    d.swap(out);
  }
}

thread_local bool g_fixed_launch = false;
thread_local int g_blue_emax = 0;  // elements per thread of a Bluestein-kernel launch

template <typename R>
void check_shape(const sa_pass_t& p) {
  auto die = [](const char* m) {
    std::fprintf(stderr, "EMU BAD LAUNCH: %s\n", m);
    std::abort();
  };
  // the fixed kernel holds 32 complex floats per thread at n = 512, 1024
  const int emax = g_blue_emax ? g_blue_emax
                               : (g_fixed_launch && sizeof(R) == 4 && (p.n == 512 || p.n == 1024) ? 32 : 16);
  if (p.T < 1 || p.C < 1 || p.C * p.T > 512) die("threads per workgroup");
  if ((int64_t)p.n > (int64_t)emax * p.T) die("n > EMAX * T");
  if (!g_fixed_launch) {
    if (p.emax != 8 && p.emax != 16) die("elements per thread");
    if (p.emax == 8 && (!p.blu_h || sizeof(R) != 4))
      die("8 elements per thread outside the complex-float Bluestein row kernel");
    if ((int64_t)p.n > (int64_t)p.emax * p.T) die("n > emax * T");
  }
  for (int s = 0; s < p.nstages; ++s) {
    const int rad = p.radix[s];
    // butterflies the generic kernel has (fft_device.hpp, pass_tile.hpp)
    static const int kRad[] = {2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 14, 15, 16, 18, 20, 21, 24, 25, 28};
    if (std::find(std::begin(kRad), std::end(kRad), rad) == std::end(kRad)) die("radix without a kernel case");
    if (sizeof(R) == 4 && rad > 21) die("radix above 21 in a complex-float pass");
    const int maxb = (emax + rad - 1) / rad;
    if (!g_fixed_launch && (p.n / rad + p.T - 1) / p.T > maxb)
      die("butterflies per thread exceed register budget");
  }
  if (p.lds_pad != 0 && p.lds_pad != 1) die("lds_pad");
  if ((p.blu_h || g_fixed_launch || g_blue_emax) && !p.lds_pad) die("unpadded rows outside the generic pass");
  if (p.ldsstride < (p.lds_pad ? p.n + (p.n - 1) / 16 : p.n)) die("lds stride");
  if (p.ldsstride % 2 == 0) die("even lds stride");
  if (p.reg_first || p.reg_last) {
    if (p.n != emax * p.T || p.nstages < 2 || p.blu_h) die("register stage on an unsuitable pass");
    auto p2 = [](int r) { return r == 2 || r == 4 || r == 8 || r == 16; };
    if (p.reg_first && !p2(p.radix[0])) die("reg_first radix");
    if (p.reg_last && !p2(p.radix[p.nstages - 1])) die("reg_last radix");
    const int id_kfast = p.reg_first ? p.ld_kfast : p.st_kfast;
    if (p.reg_last && id_kfast != p.st_kfast) die("reg_last with a different store mapping");
  }
  if (((int64_t)p.C * p.ldsstride + p.tw_lds) * (int64_t)sizeof(C<R>) > 160 * 1024) die("lds bytes");
  if (p.tw_lds < 0) die("tw_lds");
  if (p.tw_lds > 0) {
    int64_t need = 0;  // the staged table must cover every twiddle a stage reads
    for (int s = 1; s < p.nstages; ++s)
      need = std::max<int64_t>(need, (int64_t)p.twoff[s] + (int64_t)p.ns[s] * (p.radix[s] - 1));
    if (p.tw_lds < need) die("tw_lds below the stage table");
    chk("tw", p.tw, p.tw_lds - 1, sizeof(C<R>));
  }
  if (p.batch >= (int64_t(1) << 31)) die("batch");
  if (p.ptw_level > 2 || p.ld_level > 2 || p.st_level > 2 || p.ld_level2 > 2 || p.st_level2 > 2)
    die("level");
}

template <typename R>
void run_pass(const sa_pass_t& p, int sign) {
  check_shape<R>(p);
  const int n = p.n;
  const C<R>* in = reinterpret_cast<const C<R>*>(p.in);
  C<R>* outp = reinterpret_cast<C<R>*>(p.out);
  const C<R>* pre = reinterpret_cast<const C<R>*>(p.pre);
  const C<R>* post = reinterpret_cast<const C<R>*>(p.post);
  // Like the GPU, all lines are loaded before any is stored only within a
  // workgroup; emulate workgroup granularity so aliasing bugs show up.
  const int64_t Cw = p.C;
  for (int64_t b0 = 0; b0 < p.batch; b0 += Cw) {
    const int64_t bend = std::min<int64_t>(p.batch, b0 + Cw);
    std::vector<std::vector<C<R>>> lines;
    std::vector<int64_t> ooff;
    std::vector<std::array<int64_t, 3>> digs;
    for (int64_t b = b0; b < bend; ++b) {
      const int64_t b2 = b % p.n2, q = b / p.n2, b1 = q % p.n1, bb0 = q / p.n1;
      const int64_t ioff = bb0 * p.is0 + b1 * p.is1 + b2 * p.is2;
      ooff.push_back(bb0 * p.os0 + b1 * p.os1 + b2 * p.os2);
      std::array<int64_t, 3> dg = {bb0, b1, b2};
      digs.push_back(dg);
      std::vector<C<R>> d(n);
      for (int k = 0; k < n; ++k) {
        const int64_t m = (int64_t)k * p.ld_kmul + (p.ld_level >= 0 ? dg[p.ld_level] * p.ld_dmul : 0) +
                          (p.ld_level2 >= 0 ? dg[p.ld_level2] * p.ld_dmul2 : 0);
        C<R> v = 0;
        if (m < p.ld_limit) {
          chk("in", in, ioff + (int64_t)k * p.iks, sizeof(C<R>));
          v = in[ioff + (int64_t)k * p.iks];
          if (pre) {
            chk("pre", pre, m, sizeof(C<R>));
            v *= pre[m];
          }
        }
        d[k] = v;
      }
      if (!p.blu_h) {
        stages<R>(p, d, sign);
      } else {
        stages<R>(p, d, -1);
        const C<R>* h = reinterpret_cast<const C<R>*>(p.blu_h);
        const int64_t hoff = p.blu_level >= 0 ? dg[p.blu_level] * p.blu_hstride : 0;
        for (int k = 0; k < n; ++k) {
          chk("blu_h", h, hoff + k, sizeof(C<R>));
          d[k] *= h[hoff + k];
        }
        stages<R>(p, d, +1);
      }
      lines.push_back(std::move(d));
    }
    for (int64_t b = b0; b < bend; ++b) {
      const auto& d = lines[b - b0];
      const auto& dg = digs[b - b0];
      for (int k = 0; k < n; ++k) {
        C<R> v = d[k];
        if (p.ptw_level >= 0) {
          const uint64_t e = ((uint64_t)dg[p.ptw_level] * (uint64_t)k) % (uint64_t)p.ptw_N;
          const C<R>* lo = reinterpret_cast<const C<R>*>(p.ptw_lo);
          const C<R>* hi = reinterpret_cast<const C<R>*>(p.ptw_hi);
          chk("ptw_hi", hi, (int64_t)(e >> p.ptw_shift), sizeof(C<R>));
          chk("ptw_lo", lo, (int64_t)(e & ((1ull << p.ptw_shift) - 1)), sizeof(C<R>));
          const C<R> w = hi[e >> p.ptw_shift] * lo[e & ((1ull << p.ptw_shift) - 1)];
          v *= conj_if(p.ptw_conj ? std::conj(w) : w, sign);
        } else {
          const int64_t m = (int64_t)k * p.st_kmul + (p.st_level >= 0 ? dg[p.st_level] * p.st_dmul : 0) +
                            (p.st_level2 >= 0 ? dg[p.st_level2] * p.st_dmul2 : 0);
          if (m >= p.st_limit) continue;
          if (post) {
            chk("post", post, m, sizeof(C<R>));
            v *= post[m];
          }
        }
        chk("out", outp, ooff[b - b0] + (int64_t)k * p.oks, sizeof(C<R>));
        outp[ooff[b - b0] + (int64_t)k * p.oks] = v;
      }
    }
  }
}

}  // namespace

extern "C" int sa_max_lds_bytes(void) { return 160 * 1024; }

// The allocation-free execute's device array, as one registered host block
// of the same size (so every access is bounds-checked against it).
extern "C" int sa_noalloc_scratch(void** ptr, size_t* bytes) {
  static const size_t kBytes = size_t(64) << 20;
  static void* block = [] {
    void* p = std::malloc(kBytes);
    std::lock_guard<std::mutex> lk(allocs_mu());
    allocs()[(uintptr_t)p] = kBytes;
    return p;
  }();
  *ptr = block;
  *bytes = kBytes;
  return 0;
}

extern "C" int sa_launch_pass(const sa_pass_t* p, int prec, int sign, void*) {
  if (prec == 0)
    run_pass<float>(*p, sign);
  else
    run_pass<double>(*p, sign);
  return 0;
}

extern "C" int sa_fixed_supported(int n) {
  return n == 32 || n == 64 || n == 128 || n == 256 || n == 512 || n == 1024 || n == 2048 ||
         n == 4096 || n == 8192;
}

extern "C" int sa_fixed_occupancy(int, int, int, int, size_t) { return 1; }
// LDS-limited workgroups per CU (the registers are not modelled here).
extern "C" int sa_pass_occupancy(const sa_pass_t* p, int prec, int twl) {
  const int64_t lds = ((int64_t)p->C * p->ldsstride + (twl ? p->tw_lds : 0)) * (prec ? 16 : 8);
  return lds > 160 * 1024 ? 0 : (int)((160 * 1024) / std::max<int64_t>(lds, 1));
}
extern "C" int sa_fixed_elems(int n, int prec) { return (prec == 0 && (n == 512 || n == 1024)) ? 32 : 16; }
extern "C" int sa_fixed_max_threads(int) { return 512; }

// The fixed kernel has the same semantics; check its own launch shape first.
extern "C" int sa_launch_fixed(const sa_pass_t* p, int prec, int sign, int grid, void*) {
  auto die = [](const char* m) {
    std::fprintf(stderr, "EMU BAD FIXED LAUNCH: %s\n", m);
    std::abort();
  };
  if (!sa_fixed_supported(p->n)) die("length");
  if (p->T * sa_fixed_elems(p->n, prec) != p->n) die("T");
  if (p->C != (1 << p->logC)) die("C");
  if (p->C * p->T > sa_fixed_max_threads(prec)) die("threads");
  if (((int64_t)p->C * p->ldsstride + p->tw_lds) * (prec ? 16 : 8) > 160 * 1024) die("lds");
  if (p->tw_lds < 0) die("tw_lds");
  if (p->tw_lds > 0) {
    // the staged table must cover every twiddle a stage reads
    int64_t need = 0;
    for (int s = 1; s < p->nstages; ++s)
      need = std::max<int64_t>(need, (int64_t)p->twoff[s] + (int64_t)p->ns[s] * (p->radix[s] - 1));
    if (p->tw_lds < need) die("tw_lds below the stage table");
    chk("tw", p->tw, p->tw_lds - 1, prec ? 16 : 8);
  }
  if (p->ldsstride < p->n + (p->n - 1) / 16) die("stride");
  if (!p->lds_pad) die("fixed kernel rows are padded");
  if (grid < 1) die("grid");
  if (p->xcd_order < 0) die("xcd_order");
  if (p->xcd_order >= 2) {  // rows of xcd_order tiles: the walk is a bijection only if they tile the pass
    const int64_t tiles = (p->batch + p->C - 1) / p->C;
    if (tiles % p->xcd_order != 0) die("tile rows do not divide the tile count");
  }
  if (p->pre || p->post || p->blu_h || p->ld_level >= 0 || p->st_level >= 0) die("hooks");
  {
    // the kernel's compile-time radix sequence (fft_fixed.hip Radices<N>)
    static const std::map<int, std::vector<int>> plans = {
        {32, {8, 4}},       {64, {8, 8}},       {128, {16, 8}},     {256, {16, 16}},
        {512, {8, 8, 8}},   {1024, {16, 8, 8}}, {2048, {16, 16, 8}}, {4096, {16, 16, 16}},
        {8192, {16, 16, 8, 4}}};
    const auto& pl = plans.at(p->n);
    if ((int)pl.size() != p->nstages) die("stage count differs from the kernel's plan");
    for (int s = 0; s < p->nstages; ++s)
      if (pl[s] != p->radix[s]) die("radix differs from the kernel's plan");
  }
  sa_pass_t q = *p;
  q.reg_first = q.reg_last = 0;
  g_fixed_launch = true;
  const int rc = sa_launch_pass(&q, prec, sign, nullptr);
  g_fixed_launch = false;
  return rc;
}

extern "C" int sa_blue_elems(int n, int prec, int kind, int lines) {
  return (kind != SA_BLUE_ROWS && prec == 0 && n == 512 && lines > 16) ? 32 : 16;
}

extern "C" int sa_blue_supported(int n, int kind) {
  if (kind == SA_BLUE_ROWS) return n == 1024 || n == 2048 || n == 4096 || n == 8192;
  if (kind == SA_BLUE_COLS_IN || kind == SA_BLUE_COLS_OUT)
    return n == 32 || n == 64 || n == 128 || n == 256 || n == 512;
  return 0;
}

// The three-pass Bluestein kernels (fft_blue.hip): the generic semantics,
// after the checks of what those kernels assume (grid = tiles, fixed radix
// plan, lane mapping per kind, hooks per kind).
extern "C" int sa_launch_blue(const sa_pass_t* p, int prec, int kind, void*) {
  auto die = [](const char* m) {
    std::fprintf(stderr, "EMU BAD BLUESTEIN LAUNCH: %s\n", m);
    std::abort();
  };
  if (!sa_blue_supported(p->n, kind)) die("length");
  const bool row8 = kind == SA_BLUE_ROWS && p->n <= 4096 && p->emax == 8;  // the 8-element row variant
  if (p->emax != sa_blue_elems(p->n, prec, kind, p->C) && !row8) die("elements per thread");
  if (p->T * p->emax != p->n) die("T");
  if (p->C != (1 << p->logC)) die("C");
  if (p->C * p->T > 512) die("threads");
  if (((int64_t)p->C * p->ldsstride + p->tw_lds) * (prec ? 16 : 8) > 160 * 1024) die("lds");
  if (p->ldsstride < p->n + (p->n - 1) / 16) die("stride");
  if (p->tw_lds > 0) {
    int64_t need = 0;
    for (int s = 1; s < p->nstages; ++s)
      need = std::max<int64_t>(need, (int64_t)p->twoff[s] + (int64_t)p->ns[s] * (p->radix[s] - 1));
    if (p->tw_lds < need) die("tw_lds below the stage table");
    chk("tw", p->tw, p->tw_lds - 1, prec ? 16 : 8);
  }
  {
    static const std::map<int, std::vector<int>> plans = {
        {32, {8, 4}},       {64, {8, 8}},       {128, {16, 8}},     {256, {16, 16}},
        {512, {8, 8, 8}},   {1024, {16, 8, 8}}, {2048, {16, 16, 8}}, {4096, {16, 16, 16}},
        {8192, {16, 16, 8, 4}}};
    static const std::map<int, std::vector<int>> plans8 = {
        {1024, {8, 8, 4, 4}}, {2048, {8, 8, 8, 4}}, {4096, {8, 8, 8, 8}}};
    const auto& pl = row8 ? plans8.at(p->n) : plans.at(p->n);
    if ((int)pl.size() != p->nstages) die("stage count differs from the kernel's plan");
    for (int s = 0; s < p->nstages; ++s)
      if (pl[s] != p->radix[s]) die("radix differs from the kernel's plan");
  }
  if (kind == SA_BLUE_ROWS) {
    if (!p->ld_kfast || !p->st_kfast) die("row pass must be k-fast");
    if (!p->blu_h || p->pre || p->post || p->ptw_level < 0 || !p->ptw_conj) die("row pass hooks");
  } else {
    if (p->ld_kfast || p->st_kfast) die("column pass must be line-fast");
    if (p->blu_h || p->blu_level >= 0) die("column pass filter");
    if (kind == SA_BLUE_COLS_IN && (!p->pre || p->post || p->st_level >= 0 || p->ptw_level < 0))
      die("chirp-in hooks");
    if (kind == SA_BLUE_COLS_OUT && (p->pre || !p->post || p->ld_level >= 0 || p->ptw_level >= 0))
      die("chirp-out hooks");
  }
  sa_pass_t q = *p;
  q.reg_first = q.reg_last = 0;
  g_fixed_launch = true;
  g_blue_emax = p->emax;
  const int rc = sa_launch_pass(&q, prec, kind == SA_BLUE_COLS_OUT ? +1 : -1, nullptr);
  g_fixed_launch = false;
  g_blue_emax = 0;
  return rc;
}

extern "C" int sa_launch_scale(void* data, double f, uint64_t count, int prec, void*) {
  if (prec == 0) {
    float* x = (float*)data;
    const float ff = (float)f;
    for (uint64_t i = 0; i < 2 * count; ++i) x[i] *= ff;
  } else {
    double* x = (double*)data;
    for (uint64_t i = 0; i < 2 * count; ++i) x[i] *= f;
  }
  return 0;
}

// ---- batch launch: run the work list in order, one whole pass at the first
// item of each pass, after checking the list's dependency order ---------------
namespace {
std::set<int> g_done_pass;  // passes completed by earlier level launches
bool g_capturing = false;
std::vector<std::function<void()>>* g_capture = nullptr;
}

extern "C" int sa_batch_occupancy(int, int, size_t) { return 1; }

extern "C" int sa_launch_batch(const sa_batch_t* bp, int prec, int grid, int block, size_t lds,
                               void*) {
  const sa_batch_t b = *bp;
  auto run = [b, prec, grid, block, lds]() {
    auto die = [](const char* m) {
      std::fprintf(stderr, "EMU BAD BATCH: %s\n", m);
      std::abort();
    };
    if (grid < 1 || block < 1 || block > 512 || lds > 159 * 1024) die("launch shape");
    if (!b.err || b.wait_ticks == 0) die("missing wait bound");
    if (b.queues != 1 && (b.queues != 8 || !b.independent)) die("work queues");
    bool first_level = true;
    for (uint32_t it = 0; it < b.nitems; ++it)
      if (b.pass_dep[b.items[it].pass] >= 0) first_level = false;
    if (!b.independent || first_level) g_done_pass.clear();
    std::vector<uint32_t> seen(b.npasses, 0);
    std::vector<char> ran(b.npasses, 0);
    for (uint32_t it = 0; it < b.nitems; ++it) {
      const sa_batch_item_t w = b.items[it];
      if (w.pass >= b.npasses || w.tile >= b.pass_tiles[w.pass]) die("item out of range");
      const int dep = b.pass_dep[w.pass];
      if (!b.independent && dep >= 0 && seen[dep] != b.pass_tiles[dep])
        die("item precedes its dependency");
      if (b.independent && dep >= 0 && !g_done_pass.count(dep)) die("level launched before its dependency");
      const sa_pass_t& p = b.passes[w.pass];
      if (p.C * p.T > block) die("pass tile larger than the workgroup");
      if ((size_t)p.C * p.ldsstride * (prec ? 16 : 8) > lds) die("pass LDS larger than the launch");
      if (!ran[w.pass]) {
        sa_launch_pass(&p, prec, b.pass_sign[w.pass], nullptr);
        ran[w.pass] = 1;
      }
      ++seen[w.pass];
    }
    if (!b.independent) {
      for (uint32_t i = 0; i < b.npasses; ++i)
        if (seen[i] != b.pass_tiles[i]) die("pass not fully covered by the work list");
    } else {
      for (uint32_t i = 0; i < b.npasses; ++i)
        if (seen[i] != 0 && seen[i] != b.pass_tiles[i]) die("level launch covers part of a pass");
      for (uint32_t i = 0; i < b.npasses; ++i)
        if (seen[i]) g_done_pass.insert((int)i);
    }
  };
  if (g_capturing)
    g_capture->push_back(run);
  else
    run();
  return 0;
}

// ---- the HIP runtime calls used by the host code, on host memory ----------
extern "C" {
hipError_t hipMalloc(void** p, size_t n) {
  *p = std::malloc(n ? n : 1);
  std::lock_guard<std::mutex> lk(allocs_mu());
  if (*p) allocs()[(uintptr_t)*p] = n;
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipHostMalloc(void** p, size_t n, unsigned int) { return hipMalloc(p, n); }
hipError_t hipHostFree(void* p) { return hipFree(p); }
hipError_t hipFree(void* p) {
  std::lock_guard<std::mutex> lk(allocs_mu());
  allocs().erase((uintptr_t)p);
  std::free(p);
  return hipSuccess;
}
hipError_t hipMemset(void* d, int v, size_t n) {
  std::memset(d, v, n);
  return hipSuccess;
}
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) {
  std::memmove(d, s, n);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t) {
  if (g_capturing) {
    g_capture->push_back([d, s, n]() { std::memmove(d, s, n); });
    return hipSuccess;
  }
  std::memmove(d, s, n);
  return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned int) {
  *s = reinterpret_cast<hipStream_t>(0x1);
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) {
  g_capturing = true;
  g_capture = new std::vector<std::function<void()>>();
  return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) {
  g_capturing = false;
  *g = reinterpret_cast<hipGraph_t>(g_capture);
  g_capture = nullptr;
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, hipGraphNode_t*, char*, size_t) {
  *e = reinterpret_cast<hipGraphExec_t>(g);
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
  for (auto& f : *reinterpret_cast<std::vector<std::function<void()>>*>(e)) f();
  return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }
hipError_t hipGraphDestroy(hipGraph_t g) {
  delete reinterpret_cast<std::vector<std::function<void()>>*>(g);
  return hipSuccess;
}
hipError_t hipEventCreate(hipEvent_t* e) {
  *e = reinterpret_cast<hipEvent_t>(0x1);
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) {
  *ms = 1.0f;
  return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
// Emulated launches complete before they return, so every event is done.
hipError_t hipEventQuery(hipEvent_t) { return hipSuccess; }
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
const char* hipGetErrorString(hipError_t) { return "emulated"; }
hipError_t hipGetDevice(int* d) {
  *d = 0;
  return hipSuccess;
}
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t a, int) {
  *v = (a == hipDeviceAttributeMaxSharedMemoryPerBlock) ? 160 * 1024 : 0;
  return hipSuccess;
}
// No device to name: configuration objects get deviceName NULL.
hipError_t hipGetDeviceCount(int* n) {
  *n = 0;
  return hipSuccess;
}
hipError_t hipGetDeviceProperties(hipDeviceProp_t*, int) { return hipErrorNoDevice; }
hipError_t hipGetLastError() { return hipSuccess; }
}
