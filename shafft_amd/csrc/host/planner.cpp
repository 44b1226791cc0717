#include "planner.hpp"

#include <algorithm>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <sstream>

#include "fastdiv.hpp"
#include "tables.hpp"

namespace sa {

// ---------------------------------------------------------------------------
DevBuf::DevBuf(size_t n) : bytes(n) {
  if (n) hip_check(hipMalloc(&p, n), "hipMalloc(table/scratch)");
}
DevBuf::~DevBuf() {
  if (p) (void)hipFree(p);
}
DevBuf& DevBuf::operator=(DevBuf&& o) noexcept {
  if (this != &o) {
    if (p) (void)hipFree(p);
    p = o.p;
    bytes = o.bytes;
    o.p = nullptr;
    o.bytes = 0;
  }
  return *this;
}

// ---------------------------------------------------------------------------
// Device queries are made once per process (thread-safe static init): plans
// may be built on several host threads at once.
static int lds_budget_bytes() {
  static const int cached = [] {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess &&
        v > 0)
      return std::min(v, sa_max_lds_bytes());
    return 64 * 1024;
  }();
  return cached;
}

// Lines up to this length run as one LDS-resident pass.  A line must fit
// the LDS of one workgroup with room for a second one per CU.
// SHAFFT_AMD_LDS_CAP (elements, >= 8) lowers the cap: used to exercise the
// multi-level four-step and Bluestein plans on small inputs.
int64_t lds_cap(int prec) {
  const int64_t hw = prec ? 4096 : 8192;
  if (const char* e = std::getenv("SHAFFT_AMD_LDS_CAP")) {
    const long long v = std::atoll(e);
    if (v >= 8 && v < hw) return v;
  }
  return hw;
}

// Longest line for a pass whose elements are strided in memory.  Such a pass
// reads (or writes) C adjacent lines per element index, so its HBM segments
// are C * element-size bytes; with 16 elements per thread and 512 threads
// per workgroup, C >= 16 needs n <= 512.  Lines up to 1024 still run as one
// pass with fewer, narrower lines: splitting them costs a whole extra pass,
// and one pass at 64-byte segments beats two at 128 (serving-stream jobs
// with a 525..675-point strided axis or a 625-point four-step factor:
// 4 -> 3 passes, 12-20 % less time per job).  Longer ones are split
// (four-step).
int64_t lf_cap(int prec) {
  int64_t c = 1024;
  if (const char* e = std::getenv("SHAFFT_AMD_LF_CAP")) {
    const long long v = std::atoll(e);
    if (v >= 8) c = v;
  }
  return std::min(c, lds_cap(prec));
}

// A strided line of length n stays one pass when it is under lf_cap, or
// when the fixed-length kernel runs it with 16-line tiles (1024-point
// complex float: 32 elements per thread, 512 threads, 139 KB of LDS).
static bool strided_single(int64_t n, int prec) {
  if (n <= lf_cap(prec)) return true;
  if (std::getenv("SHAFFT_AMD_LF_CAP") || n > lds_cap(prec)) return false;
  if (!sa_fixed_supported((int)n)) return false;
  const int64_t T = n / sa_fixed_elems((int)n, prec);
  const int64_t stride = n + (n - 1) / 16 + 1;
  return 16 * T <= sa_fixed_max_threads(prec) && 16 * stride * (prec ? 16 : 8) <= 160 * 1024;
}

bool is_smooth7(int64_t n) {
  if (n < 1) return false;
  for (int p : {2, 3, 5, 7})
    while (n % p == 0) n /= p;
  return n == 1;
}

static int env_int(const char* name, int dflt);

static std::vector<int> radix_plan_greedy(int64_t n) {
  std::vector<int> r;
  int a = 0;
  while (n % 2 == 0) {
    n /= 2;
    ++a;
  }
  // Powers of two: radix-16 stages, the remainder as one radix 8/4/2 stage;
  // 2^(4q+1) is taken as 16^(q-1) * 8 * 4 rather than 16^q * 2.
  int q = a / 4, rem = a % 4;
  if (rem == 1 && q >= 1) {
    for (int i = 0; i < q - 1; ++i) r.push_back(16);
    r.push_back(8);
    r.push_back(4);
  } else {
    for (int i = 0; i < q; ++i) r.push_back(16);
    if (rem == 3) r.push_back(8);
    if (rem == 2) r.push_back(4);
    if (rem == 1) r.push_back(2);
  }
  int e3 = 0, e5 = 0, e7 = 0;
  for (; n % 3 == 0; n /= 3) ++e3;
  for (; n % 5 == 0; n /= 5) ++e5;
  for (; n % 7 == 0; n /= 7) ++e7;
  if (n != 1) fail(ERR_INTERNAL, "radix_plan: length is not 7-smooth");
  // Fewer stages = fewer LDS round trips and barriers: threes pair into
  // radix 9, and a lone 3 or 5 merges with the trailing radix-2/4 stage
  // into the prime-factor butterflies 6, 12 or 10.
  const bool tail2 = !r.empty() && r.back() == 2, tail4 = !r.empty() && r.back() == 4;
  int e9 = e3 / 2;
  e3 %= 2;
  int r6 = 0, r12 = 0, r10 = 0;
  if (e3 == 1 && tail2) {
    r.pop_back();
    r6 = 1;
    e3 = 0;
  } else if (e3 == 1 && tail4) {
    r.pop_back();
    r12 = 1;
    e3 = 0;
  } else if (e5 > 0 && tail2) {
    r.pop_back();
    r10 = 1;
    --e5;
  }
  if (r12) r.push_back(12);
  if (r10) r.push_back(10);
  for (int i = 0; i < e9; ++i) r.push_back(9);
  if (r6) r.push_back(6);
  for (int i = 0; i < e7; ++i) r.push_back(7);
  for (int i = 0; i < e5; ++i) r.push_back(5);
  if (e3) r.push_back(3);
  return r;
}

// Radix sequence with the fewest, cheapest stages.  Every LDS stage costs
// about the same per element (~20-30 VALU instructions: addresses, twiddle,
// butterfly share, barrier), so the plan minimises an estimate of
// sum(stage cost) over factorisations of n into the kernel's butterflies:
//   cost(R) = (butterfly ops + 4 (R-1) twiddle + 8 R address + 10) / R
//             x rounds factor (per-line numbering runs ceil(16/R) rounds for
//               16/R butterflies per thread; radices above 16 are numbered
//               densely over the tile, one round).
// Powers of two go first (largest first), then the rest, largest first.
std::vector<int> radix_plan(int64_t n, int maxrad) {
  if (!env_int("SHAFFT_AMD_WIDE", 1)) return radix_plan_greedy(n);
  int e[4] = {0, 0, 0, 0};
  int64_t m = n;
  const int pr[4] = {2, 3, 5, 7};
  for (int i = 0; i < 4; ++i)
    for (; m % pr[i] == 0; m /= pr[i]) ++e[i];
  if (m != 1) fail(ERR_INTERNAL, "radix_plan: length is not 7-smooth");
  struct Rad {
    int r, e2, e3, e5, e7;
    double ops;  // butterfly arithmetic, instructions per butterfly
  };
  static const Rad rads[] = {
      {2, 1, 0, 0, 0, 4},     {3, 0, 1, 0, 0, 12},    {4, 2, 0, 0, 0, 16},   {5, 0, 0, 1, 0, 34},
      {6, 1, 1, 0, 0, 36},    {7, 0, 0, 0, 1, 70},    {8, 3, 0, 0, 0, 52},   {9, 0, 2, 0, 0, 96},
      {10, 1, 0, 1, 0, 88},   {12, 2, 1, 0, 0, 96},   {14, 1, 0, 0, 1, 168}, {15, 0, 1, 1, 0, 162},
      {16, 4, 0, 0, 0, 144},  {18, 1, 2, 0, 0, 228},  {20, 2, 0, 1, 0, 216}, {21, 0, 1, 0, 1, 294},
      {24, 3, 1, 0, 0, 252},  {25, 0, 0, 2, 0, 436},  {28, 2, 0, 0, 1, 392}};
  maxrad = std::min(maxrad, env_int("SHAFFT_AMD_MAXRAD", maxrad));  // can only lower it
  auto cost = [](const Rad& d) {
    const int R = d.r;
    const double per = (d.ops + 4.0 * (R - 1) + 8.0 * R + 10.0) / R;
    const double rounds = R > 16 ? 1.0 : (double)((16 + R - 1) / R) * R / 16.0;
    return per * rounds;
  };
  // DP over exponent vectors (small: n <= 2^31).
  const int A = e[0] + 1, B = e[1] + 1, C5 = e[2] + 1, D = e[3] + 1;
  auto id = [&](int a, int b, int c, int d) { return ((a * B + b) * C5 + c) * D + d; };
  const int S = A * B * C5 * D;
  std::vector<double> best(S, 1e300);
  std::vector<int> choice(S, -1);
  best[0] = 0.0;
  for (int a = 0; a < A; ++a)
    for (int b = 0; b < B; ++b)
      for (int c = 0; c < C5; ++c)
        for (int d = 0; d < D; ++d) {
          const int k = id(a, b, c, d);
          if (k == 0) continue;
          for (int j = 0; j < (int)(sizeof(rads) / sizeof(rads[0])); ++j) {
            const Rad& q = rads[j];
            if (q.r > maxrad || q.e2 > a || q.e3 > b || q.e5 > c || q.e7 > d) continue;
            const int prev = id(a - q.e2, b - q.e3, c - q.e5, d - q.e7);
            const double v = best[prev] + cost(q);
            if (v < best[k] - 1e-9) {
              best[k] = v;
              choice[k] = j;
            }
          }
        }
  std::vector<int> r;
  for (int a = e[0], b = e[1], c = e[2], d = e[3]; id(a, b, c, d) != 0;) {
    const Rad& q = rads[choice[id(a, b, c, d)]];
    r.push_back(q.r);
    a -= q.e2;
    b -= q.e3;
    c -= q.e5;
    d -= q.e7;
  }
  std::stable_sort(r.begin(), r.end(), [](int x, int y) {
    const bool px = (x & (x - 1)) == 0, py = (y & (y - 1)) == 0;
    if (px != py) return px;
    return x > y;
  });
  return r;
}

// Radix sequence the fixed-length kernels are compiled with
// (fixed_stages.hpp, Plan<N, E>); the twiddle table must follow the same
// sequence.  With 8 elements per thread (Bluestein rows) every radix
// divides 8.
static std::vector<int> fixed_radix_plan(int64_t n, int elems = 16) {
  if (elems == 8) {
    switch (n) {
      case 1024: return {8, 8, 4, 4};
      case 2048: return {8, 8, 8, 4};
      case 4096: return {8, 8, 8, 8};
      default: return {};
    }
  }
  switch (n) {
    case 32: return {8, 4};
    case 64: return {8, 8};
    case 128: return {16, 8};
    case 256: return {16, 16};
    case 512: return {8, 8, 8};
    case 1024: return {16, 8, 8};
    case 2048: return {16, 16, 8};
    case 4096: return {16, 16, 16};
    case 8192: return {16, 16, 8, 4};
    default: return {};
  }
}

static int64_t next_pow2(int64_t x) {
  int64_t m = 1;
  while (m < x) m <<= 1;
  return m;
}

// ---------------------------------------------------------------------------
const void* PassProgram::upload(const std::vector<cld>& t) {
  const size_t n = t.size();
  DevBuf b(n * (prec_ ? 16 : 8));
  if (prec_) {
    std::vector<double> h(2 * n);
    for (size_t i = 0; i < n; ++i) {
      h[2 * i] = (double)t[i].real();
      h[2 * i + 1] = (double)t[i].imag();
    }
    hip_check(hipMemcpy(b.p, h.data(), b.bytes, hipMemcpyHostToDevice), "hipMemcpy(table)");
  } else {
    std::vector<float> h(2 * n);
    for (size_t i = 0; i < n; ++i) {
      h[2 * i] = (float)t[i].real();
      h[2 * i + 1] = (float)t[i].imag();
    }
    hip_check(hipMemcpy(b.p, h.data(), b.bytes, hipMemcpyHostToDevice), "hipMemcpy(table)");
  }
  const void* p = b.p;
  tables_.push_back(std::move(b));
  return p;
}

const void* PassProgram::stage_table(const std::vector<int>& radices, std::vector<int>& offsets) {
  // One table per radix sequence, shared by every pass that uses it.
  auto it = stage_cache_.find(radices);
  if (it != stage_cache_.end()) {
    offsets = it->second.second;
    return it->second.first;
  }
  const void* t = upload(stage_twiddles(radices, offsets));
  stage_cache_[radices] = {t, offsets};
  return t;
}

// Merge adjacent digits that address memory as one digit on both sides and
// drop unit digits, never touching the digits listed in `pinned`.
static std::vector<Digit> merge_digits(const std::vector<Digit>& in, std::vector<int>& pinned) {
  std::vector<Digit> d;
  std::vector<int> map(in.size(), -1);
  for (size_t j = 0; j < in.size(); ++j) {
    const bool pin = std::find(pinned.begin(), pinned.end(), (int)j) != pinned.end();
    if (in[j].cnt == 1 && !pin) continue;
    if (!d.empty() && !pin) {
      // previous kept digit
      int prev = -1;
      for (int i = (int)j - 1; i >= 0; --i)
        if (map[i] >= 0) {
          prev = i;
          break;
        }
      const bool prev_pin =
          prev >= 0 && std::find(pinned.begin(), pinned.end(), prev) != pinned.end();
      Digit& o = d.back();
      if (!prev_pin && o.is == in[j].cnt * in[j].is && o.os == in[j].cnt * in[j].os) {
        o.cnt *= in[j].cnt;
        o.is = in[j].is;
        o.os = in[j].os;
        map[j] = (int)d.size() - 1;
        continue;
      }
    }
    d.push_back(in[j]);
    map[j] = (int)d.size() - 1;
  }
  for (int& p : pinned) p = map[p];
  return d;
}

// Tuning knobs read at plan time (used by the schedule sweep and tests):
//   SHAFFT_AMD_LF_LINES  lines per workgroup for passes with a line-fast side
//   SHAFFT_AMD_NT        1: non-temporal data loads/stores
//   SHAFFT_AMD_REGSTAGE  0: disable the register-resident first/last stage
static int env_int(const char* name, int dflt) {
  const char* e = std::getenv(name);
  return e && *e ? std::atoi(e) : dflt;
}

static bool pow2_le16(int r) { return r == 2 || r == 4 || r == 8 || r == 16; }

// LDS bank cycles of the first wave of a tile of the generic pass (the
// other waves repeat its lane patterns on other lines and elements), summed
// over every LDS access instruction, for padded (k -> k + k/16) or unpadded
// rows of `stride` elements.  Model: `banks` banks of 4 bytes; a wave's
// access of `w` dwords per lane is served in groups of banks/w lanes, each
// group taking as many cycles as the most dwords that fall on one bank
// (lanes repeating the previous lane's element, as clamped butterflies do,
// count once).  Follows the kernel's lane mappings (pass_tile.hpp): the
// load's and store's line slots, the butterfly identity of the stages,
// clamped butterflies in running slots, one butterfly per thread for
// radices > 16, and the register-resident first and last stages.
static int64_t lds_bank_cycles(const sa_pass_t& p, int elem_bytes, bool pad, int stride, int banks) {
  const int n = p.n, T = p.T, C = p.C, E = p.emax;
  const int nthr = std::min(C * T, 64);
  const int w = elem_bytes / 4;
  const int group = std::max(1, banks / w);
  auto lpk = [pad](int k) { return pad ? k + (k >> 4) : k; };
  auto line_of = [&](int t, bool kfast) { return kfast ? t / T : t % C; };
  auto elem_of = [&](int t, bool kfast) { return kfast ? t % T : t / C; };
  int64_t total = 0;
  std::vector<int> cnt(banks);
  std::vector<int64_t> addr(64);
  std::vector<char> on(64);
  // One instruction of one wave: lanes with on[l] access element addr[l].
  auto issue = [&](int lanes) {
    for (int g0 = 0; g0 < lanes; g0 += group) {
      std::fill(cnt.begin(), cnt.end(), 0);
      int worst = 0;
      for (int l = g0; l < std::min(lanes, g0 + group); ++l) {
        if (!on[l]) continue;
        if (l > g0 && on[l - 1] && addr[l - 1] == addr[l]) continue;  // same element: one access
        for (int d = 0; d < w; ++d) {
          const int b = (int)((addr[l] * w + d) % banks);
          worst = std::max(worst, ++cnt[b]);
        }
      }
      total += worst;
    }
  };
  const bool id_kfast = p.reg_first ? p.ld_kfast != 0 : p.st_kfast != 0;
  for (int w0 = 0; w0 < nthr; w0 += 64) {
    const int lanes = std::min(64, nthr - w0);
    // load into LDS
    if (!p.reg_first) {
      for (int i = 0; i < E; ++i) {
        for (int l = 0; l < lanes; ++l) {
          const int t = w0 + l, k = elem_of(t, p.ld_kfast != 0) + i * T;
          on[l] = k < n;
          addr[l] = (int64_t)line_of(t, p.ld_kfast != 0) * stride + lpk(k);
        }
        issue(lanes);
      }
    }
    const int s0 = p.reg_first ? 1 : 0;
    const int s1 = p.reg_last ? p.nstages - 1 : p.nstages;
    if (p.reg_first) {
      const int R = p.radix[0];
      for (int i = 0; i < 16 / R; ++i)
        for (int r = 0; r < R; ++r) {
          for (int l = 0; l < lanes; ++l) {
            const int t = w0 + l;
            on[l] = 1;
            addr[l] = (int64_t)line_of(t, id_kfast) * stride + lpk((elem_of(t, id_kfast) + i * T) * R + r);
          }
          issue(lanes);
        }
    }
    for (int s = s0; s < s1; ++s) {
      const int R = p.radix[s], Ns = p.ns[s], nb = n / R;
      if (R > 16) {
        for (int phase = 0; phase < 2; ++phase)
          for (int r = 0; r < R; ++r) {
            for (int l = 0; l < lanes; ++l) {
              const int t = w0 + l;
              on[l] = t < C * nb;
              const int c = t / nb, jb = t % nb;
              const int q = jb / Ns, rem = jb - q * Ns;
              const int k = phase == 0 ? jb + r * nb : q * Ns * R + rem + r * Ns;
              addr[l] = (int64_t)c * stride + lpk(k);
            }
            issue(lanes);
          }
        continue;
      }
      const int maxb = (E + R - 1) / R;
      for (int i = 0; i < maxb; ++i) {
        bool run = false;
        for (int l = 0; l < lanes && !run; ++l) run = elem_of(w0 + l, id_kfast) + i * T < nb;
        if (!run) continue;
        for (int phase = 0; phase < 2; ++phase)
          for (int r = 0; r < R; ++r) {
            for (int l = 0; l < lanes; ++l) {
              const int t = w0 + l, tl = elem_of(t, id_kfast);
              const int jb = std::min(tl + i * T, nb - 1);
              const int q = jb / Ns, rem = jb - q * Ns;
              on[l] = phase == 0 || tl + i * T < nb;
              const int k = phase == 0 ? jb + r * nb : q * Ns * R + rem + r * Ns;
              addr[l] = (int64_t)line_of(t, id_kfast) * stride + lpk(k);
            }
            issue(lanes);
          }
      }
    }
    if (p.reg_last) {
      const int R = p.radix[p.nstages - 1], nb = n / R;
      for (int i = 0; i < 16 / R; ++i)
        for (int r = 0; r < R; ++r) {
          for (int l = 0; l < lanes; ++l) {
            const int t = w0 + l;
            on[l] = 1;
            addr[l] = (int64_t)line_of(t, id_kfast) * stride + lpk(elem_of(t, id_kfast) + i * T + r * nb);
          }
          issue(lanes);
        }
    } else {
      for (int i = 0; i < E; ++i) {
        for (int l = 0; l < lanes; ++l) {
          const int t = w0 + l;
          on[l] = 1;
          addr[l] = (int64_t)line_of(t, p.st_kfast != 0) * stride +
                    lpk(std::min(elem_of(t, p.st_kfast != 0) + i * T, n - 1));
        }
        issue(lanes);
      }
    }
  }
  return total;
}

// Padded or unpadded LDS rows for a generic pass (p.C, p.T, lane mappings
// and register-stage flags already set).  Unpadded rows save three VALU
// instructions per LDS access (the generic pass is VALU-issue bound) and
// are taken when the bank model gives them at most `slack` times the padded
// rows' cycles, on both a 64-bank and a 32-bank LDS.
//   SHAFFT_AMD_LDS_PAD: 1 always padded, 0 always unpadded, -1 (default) model.
static bool choose_lds_pad(const sa_pass_t& p, int elem_bytes, int padded_stride, int plain_stride) {
  const int knob = env_int("SHAFFT_AMD_LDS_PAD", -1);
  if (knob >= 0) return knob != 0;
  const int slack_pct = env_int("SHAFFT_AMD_LDS_SLACK_PCT", 110);
  // Decisions are cached per launch shape (plans of one server repeat them).
  std::vector<int64_t> key = {p.n, p.T, p.C, p.emax, p.nstages, p.reg_first, p.reg_last, p.ld_kfast,
                              p.st_kfast, elem_bytes, padded_stride, plain_stride, slack_pct};
  for (int s = 0; s < p.nstages; ++s) key.push_back(p.radix[s]);
  static std::mutex mu;
  static std::map<std::vector<int64_t>, bool> cache;
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
  }
  bool pad = false;
  for (int banks : {64, 32}) {
    const int64_t cp = lds_bank_cycles(p, elem_bytes, true, padded_stride, banks);
    const int64_t cu = lds_bank_cycles(p, elem_bytes, false, plain_stride, banks);
    if ((double)cu * 100.0 > (double)slack_pct * (double)cp) {
      pad = true;
      break;
    }
  }
  std::lock_guard<std::mutex> lk(mu);
  cache.emplace(std::move(key), pad);
  return pad;
}

// Fill launch shape: threads per line, lines per workgroup, lane mappings,
// LDS stride and the register-resident stage flags.
static void configure_launch(sa_pass_t& p, int elem_bytes, bool bluestein) {
  const int n = p.n;
  // 8 elements per thread for long complex-float in-LDS Bluestein rows (the
  // middle pass of the three-pass form: few rows, 2x the threads per row;
  // 1,000,003-point FFT1D 0.062 -> 0.059 ms).  Complex double keeps 16: its
  // row kernel needs 177 VGPRs, and 512-thread rows would halve the
  // workgroups per CU (0.079 -> 0.093 ms).
  const int emax = bluestein && elem_bytes == 8 && n >= 2048 && n <= 8 * 512 &&
                           env_int("SHAFFT_AMD_BLUE_E8", 1)
                       ? 8
                       : 16;
  int T = (n + emax - 1) / emax;
  if (T < 1) T = 1;
  if (T > 512) fail(ERR_INTERNAL, "pass length exceeds the per-workgroup thread budget");
  const bool ld_lf = p.iks != 1 && p.is2 == 1 && p.n2 > 1;
  const bool st_lf = p.oks != 1 && p.os2 == 1 && p.n2 > 1;
  p.ld_kfast = ld_lf ? 0 : 1;
  p.st_kfast = st_lf ? 0 : 1;
  const bool lf = ld_lf || st_lf;
  int C = std::max(1, 256 / T);
  if (lf) {
    C = std::max(C, 256 / elem_bytes);  // 256-byte row segments when the threads allow
    const int knob = env_int("SHAFFT_AMD_LF_LINES", 0);
    if (knob > 0) C = knob;
    // line-fast lanes walk lines: keep C within the fast digit
    C = (int)std::min<int64_t>(C, p.n2);
  }
  // Tiles of at most 64 KiB of data (512 threads for complex float, 256 for
  // complex double; more only when one line needs it): several tiles per CU
  // overlap one's loads with another's stages (measured best of 128..512
  // threads over the serving stream's largest jobs, per precision).
  const int thr_dflt = 65536 / (emax * elem_bytes);
  const int max_thr = std::max(T, std::min(512, env_int("SHAFFT_AMD_WG_THREADS", thr_dflt)));
  while (C > 1 && C * T > max_thr) C /= 2;
  // A line-fast pass takes as many lines as the thread budget holds, not
  // the power of two below it: wider row segments (405-point complex float:
  // 19 lines, 152 bytes, instead of 16) and long strided lines past 512
  // points (625: 12 lines instead of 8); the serving stream's largest jobs
  // 12.5 -> 11.8 ms summed (SHAFFT_AMD_LF_CFIT=0: powers of two).
  if (lf && env_int("SHAFFT_AMD_LF_CFIT", 1))
    C = (int)std::min<int64_t>(std::max(C, max_thr / T), p.n2);
  // padded row: index k -> k + k/16, odd stride so line-fast lanes spread
  int stride = n + (n - 1) / 16;
  if (stride % 2 == 0) stride += 1;
  const int budget = lds_budget_bytes();
  while (C > 1 && (int64_t)C * stride * elem_bytes > budget) C /= 2;
  if ((int64_t)C * stride * elem_bytes > budget) fail(ERR_INTERNAL, "pass does not fit LDS");
  if (C > p.batch) C = (int)std::max<int64_t>(1, p.batch);
  p.T = T;
  p.emax = emax;
  p.C = C;
  p.ldsstride = stride;
  p.Tdiv = make_fastdiv((uint32_t)T);
  p.Cdiv = make_fastdiv((uint32_t)C);
  const bool regs_ok = !bluestein && env_int("SHAFFT_AMD_REGSTAGE", 1) != 0 && n == emax * T &&
                       p.nstages >= 2;
  // Register-resident first and last stage (one kernel variant): needs
  // radices dividing 16 at both ends and one lane mapping for load and store.
  const bool regs = regs_ok && pow2_le16(p.radix[0]) && pow2_le16(p.radix[p.nstages - 1]) &&
                    p.ld_kfast == p.st_kfast;
  p.reg_first = regs ? 1 : 0;
  p.reg_last = regs ? 1 : 0;
  // Unpadded LDS rows (odd stride, so line-fast lanes still spread) when
  // the bank model allows; Bluestein passes keep the padded rows.
  p.lds_pad = 1;
  if (!bluestein) {
    const int plain = n % 2 ? n : n + 1;
    if (!choose_lds_pad(p, elem_bytes, stride, plain)) {
      p.lds_pad = 0;
      stride = plain;
      p.ldsstride = stride;
    }
  }
  p.nt = env_int("SHAFFT_AMD_NT", 1) ? 1 : 0;
  // XCD-contiguous tile order (each XCD walks its own range of adjacent
  // tiles) for passes with a line-fast side of row segments up to 1 KiB:
  // adjacent tiles share the cache lines at their segment edges, which
  // then meet in one L2.  Measured on the serving stream's 25 largest jobs
  // per precision: median +5 % (<= 128 B segments), +24 % / +44 % (<= 512 B,
  // complex float / double), +6 % / +17 % (<= 1 KiB); passes read and
  // written k-fast, or with segments over 1 KiB, run 2-7 % faster in the
  // round-robin order.  SHAFFT_AMD_XCD_GEN: 0 never, 1 always.
  const int xcd_knob = env_int("SHAFFT_AMD_XCD_GEN", -1);
  p.xcd_order = xcd_knob >= 0 ? (xcd_knob ? 1 : 0) : (lf && C * elem_bytes <= 1024 ? 1 : 0);
  // Stage twiddles staged in LDS behind the tile when they fit.  In-LDS
  // Bluestein passes keep reading them from global memory by default
  // (SHAFFT_AMD_BTW_LDS=1 stages them: 1,000,003-point row pass 22.9 ->
  // 23.9 us complex float, 37.4 -> 37.8 us complex double).
  int64_t tw_elems = 0;
  for (int s = 1; s < p.nstages; ++s)
    tw_elems = std::max<int64_t>(tw_elems, (int64_t)p.twoff[s] + (int64_t)p.ns[s] * (p.radix[s] - 1));
  p.tw_lds = 0;
  if (env_int(bluestein ? "SHAFFT_AMD_BTW_LDS" : "SHAFFT_AMD_GTW_LDS", bluestein ? 0 : 1) &&
      tw_elems > 0 &&
      ((int64_t)C * stride + tw_elems) * elem_bytes <= budget) {
    // ... unless the table costs workgroups per CU: long single-line tiles
    // (n ~ 6000, a table of ~n elements) would drop from 2 to 1 per CU,
    // 2.9 -> 1.6 TB/s.
    p.tw_lds = (int32_t)tw_elems;
    const int prec = elem_bytes == 16 ? 1 : 0;
    if (sa_pass_occupancy(&p, prec, 1) < sa_pass_occupancy(&p, prec, 0)) p.tw_lds = 0;
  }
}

static int device_cus() {
  static const int cached = [] {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
      return v;
    return 256;
  }();
  return cached;
}

// Switch a plain power-of-two pass to the fixed-length persistent kernel:
// lines per workgroup C (power of two), LDS stride and persistent grid.
//   SHAFFT_AMD_FIXED=0 disables it; SHAFFT_AMD_FX_C / SHAFFT_AMD_FX_CK
//   override C for passes with / without a line-fast side.
static bool configure_fixed(sa_pass_t& p, int prec, int64_t& grid) {
  const int n = p.n;
  if (!env_int("SHAFFT_AMD_FIXED", 1) || !sa_fixed_supported(n)) return false;
  if (p.pre || p.post || p.blu_h || p.ld_level >= 0 || p.st_level >= 0 || p.ld_kmul != 1 ||
      p.st_kmul != 1 || p.ld_limit < n || p.st_limit < n)
    return false;
  const int esz = prec ? 16 : 8;
  const int T = n / sa_fixed_elems(n, prec);
  if (T < 1) return false;
  const bool lf = !p.ld_kfast || !p.st_kfast;
  // Fill the 512-thread workgroup (larger tiles measured faster: 512^3
  // complex128 contiguous pass 8 -> 16 lines, forward 2.51 -> 2.46 ms); a
  // line-fast side wants 256-byte row segments (32 complex float / 16
  // complex double adjacent lines), at least 16 lines.
  int C = std::max(lf ? 256 / esz : 1, 512 / T);
  const int knob = env_int(lf ? "SHAFFT_AMD_FX_C" : "SHAFFT_AMD_FX_CK", 0);
  if (knob > 0) C = knob;
  int logC = 0;
  while ((1 << (logC + 1)) <= C) ++logC;  // round down to a power of two
  int stride = n + (n - 1) / 16;
  if (stride % 2 == 0) stride += 1;
  const int budget = lds_budget_bytes();
  const int64_t maxt = sa_fixed_max_threads(prec);
  while (logC > (lf ? 4 : 0) &&
         (((int64_t)T << logC) > maxt || ((int64_t)stride << logC) * esz > budget))
    --logC;
  if (((int64_t)stride << logC) * esz > budget || ((int64_t)T << logC) > maxt) return false;
  const int nthr = T << logC;
  size_t lds = ((size_t)stride << logC) * esz;
  // Stage twiddles in LDS when they fit beside the line slots.
  int64_t tw_elems = 0;
  for (int s = 1; s < p.nstages; ++s)
    tw_elems = std::max<int64_t>(tw_elems, (int64_t)p.twoff[s] + (int64_t)p.ns[s] * (p.radix[s] - 1));
  int32_t tw_lds = 0;
  if (env_int("SHAFFT_AMD_TW_LDS", 1) && tw_elems > 0 &&
      lds + (size_t)tw_elems * esz <= (size_t)budget) {
    tw_lds = (int32_t)tw_elems;
    lds += (size_t)tw_elems * esz;
  }
  const int occ = sa_fixed_occupancy(n, prec, -1, nthr, lds);
  if (occ <= 0) return false;  // p keeps the generic kernel's launch shape
  p.tw_lds = tw_lds;
  const int64_t tiles = (p.batch + (1 << logC) - 1) >> logC;
  {
    // Tile order.  Default: passes strided by at least 2^16 elements on both
    // sides walk rows of 32 tiles down the rows first, so the tiles in
    // flight spread over many rows of each plane instead of one 64 KiB
    // window (512^3 complex128 z pass 4966 -> 5062 GB/s, FFT1D 2^27
    // complex64 first pass 4289 -> 4436; passes of smaller stride lose and
    // keep the plain order).  SHAFFT_AMD_XCD=0: plain order everywhere;
    // =1: each XCD walks its own contiguous eighth of the tiles; =2: rows
    // of SHAFFT_AMD_XCD_ROW tiles (default 32) on every pass.
    const int knob = env_int("SHAFFT_AMD_XCD", -1);
    const int64_t rlen = knob == 2 ? env_int("SHAFFT_AMD_XCD_ROW", 32) : 32;
    const bool wide = lf && !p.ld_kfast && !p.st_kfast && p.iks >= (1 << 16) && p.oks >= (1 << 16);
    const bool rows = knob == 2 || (knob < 0 && wide);
    p.xcd_order = knob == 1 ? 1 : (rows && rlen >= 2 && tiles % rlen == 0) ? (int32_t)rlen : 0;
  }
  grid = std::min<int64_t>(tiles, (int64_t)occ * device_cus());
  p.logC = logC;
  p.C = 1 << logC;
  p.T = T;
  p.ldsstride = stride;
  p.Tdiv = make_fastdiv((uint32_t)T);
  p.Cdiv = make_fastdiv((uint32_t)p.C);
  p.reg_first = p.reg_last = 0;
  p.lds_pad = 1;  // the fixed kernel's rows are padded
  return true;
}

void PassProgram::emit_single(int64_t N, std::vector<Digit> lines, int64_t iks, int64_t oks,
                              int sign, const ElemHook& ld, const ElemHook& st, const void* blu_h,
                              int ptw_level, int64_t ptw_N, int64_t out_elems) {
  // Digits referenced by hooks or the twiddle keep their identity through
  // merging; order of `lv` below is the order they are remapped in.
  int lv[5] = {ld.level, ld.level2, st.level, st.level2, ptw_level};
  std::vector<int> pin_in;
  for (int v : lv)
    if (v >= 0) pin_in.push_back(v);
  std::vector<Digit> d = merge_digits(lines, pin_in);
  if (d.size() > 3) fail(ERR_NOT_IMPL, "pass needs more than three line digits");
  {
    int idx = 0;
    for (int& v : lv)
      if (v >= 0) v = pin_in[idx++];
  }
  int& ld_level = lv[0];
  int& ld_level2 = lv[1];
  int& st_level = lv[2];
  int& st_level2 = lv[3];
  int& pt_level = lv[4];
  // The lane-fast digit of a line-fast side must be the last one (b2):
  // move a unit-stride digit there when the element stride is not unit.
  {
    int cand = -1;
    for (int j = 0; j < (int)d.size(); ++j)
      if (iks != 1 && oks != 1 && d[j].is == 1 && d[j].os == 1) cand = j;
    for (int j = 0; cand < 0 && j < (int)d.size(); ++j)
      if (oks != 1 && d[j].os == 1) cand = j;
    for (int j = 0; cand < 0 && j < (int)d.size(); ++j)
      if (iks != 1 && d[j].is == 1) cand = j;
    const int last = (int)d.size() - 1;
    if (cand >= 0 && cand != last) {
      Digit mv = d[cand];
      d.erase(d.begin() + cand);
      d.push_back(mv);
      auto remap = [&](int& lv) {
        if (lv < 0) return;
        if (lv == cand) lv = last;
        else if (lv > cand) --lv;
      };
      for (int& v : lv) remap(v);
    }
  }
  // left-pad to three digits
  const int pad = 3 - (int)d.size();
  while (d.size() < 3) d.insert(d.begin(), Digit{1, 0, 0});
  for (int& v : lv)
    if (v >= 0) v += pad;

  PassSpec ps;
  sa_pass_t& p = ps.p;
  std::memset(&p, 0, sizeof(p));
  int64_t batch = d[0].cnt * d[1].cnt * d[2].cnt;
  if (batch >= (int64_t(1) << 31) || d[1].cnt >= (int64_t(1) << 31) ||
      d[2].cnt >= (int64_t(1) << 31))
    fail(ERR_SIZE_OVERFLOW, "pass line count exceeds 2^31");
  p.batch = batch;
  p.n1 = d[1].cnt;
  p.n2 = d[2].cnt;
  p.is0 = d[0].is;
  p.is1 = d[1].is;
  p.is2 = d[2].is;
  p.os0 = d[0].os;
  p.os1 = d[1].os;
  p.os2 = d[2].os;
  p.iks = iks;
  p.oks = oks;
  p.n = (int32_t)N;
  p.n1div = make_fastdiv((uint32_t)p.n1);
  p.n2div = make_fastdiv((uint32_t)p.n2);

  // The fixed-length kernel (when eligible) dictates the radix sequence.
  const bool fixed_candidate = env_int("SHAFFT_AMD_FIXED", 1) && sa_fixed_supported((int)N) &&
                               blu_h == nullptr && !ld.active() && !st.active();
  std::vector<int> rad =
      N > 1 ? (fixed_candidate ? fixed_radix_plan(N) : radix_plan(N, prec_ ? 28 : 21))
            : std::vector<int>{};
  if (rad.size() > SA_MAX_STAGES) fail(ERR_INTERNAL, "too many stages");
  std::vector<int> offs;
  p.tw = stage_table(rad, offs);
  p.nstages = (int32_t)rad.size();
  int64_t Ns = 1;
  for (size_t s = 0; s < rad.size(); ++s) {
    p.radix[s] = rad[s];
    p.ns[s] = (int32_t)Ns;
    p.twoff[s] = offs[s];
    p.nsdiv[s] = make_fastdiv((uint32_t)Ns);
    p.nbdiv[s] = make_fastdiv((uint32_t)(N / rad[s]));
    Ns *= rad[s];
  }
  p.ptw_level = pt_level;
  if (pt_level >= 0) {
    std::vector<cld> lo, hi;
    int shift = 0;
    two_level(ptw_N, shift, lo, hi);
    p.ptw_shift = shift;
    p.ptw_N = ptw_N;
    p.ptw_lo = upload(lo);
    p.ptw_hi = upload(hi);
  }
  p.ld_kmul = ld.kmul;
  p.ld_dmul = ld.dmul;
  p.ld_level = ld_level;
  p.ld_dmul2 = ld.dmul2;
  p.ld_level2 = ld_level2;
  p.st_dmul2 = st.dmul2;
  p.st_level2 = st_level2;
  p.ld_limit = ld.limit >= 0 ? ld.limit : INT64_MAX;
  p.st_kmul = st.kmul;
  p.st_dmul = st.dmul;
  p.st_level = st_level;
  p.st_limit = st.limit >= 0 ? st.limit : INT64_MAX;
  p.pre = ld.table;
  p.post = st.table;
  p.blu_h = blu_h;
  p.blu_level = -1;
  configure_launch(p, prec_ ? 16 : 8, blu_h != nullptr);
  {
    int64_t grid = 0;
    if (configure_fixed(p, prec_, grid)) {
      ps.fixed = true;
      ps.grid = (int)grid;
    }
  }

  ps.sign = sign;
  ps.out_elems = out_elems;
  const int64_t esz = prec_ ? 16 : 8;
  const int64_t rd = batch * std::min<int64_t>(N, p.ld_limit < N ? p.ld_limit : N);
  const int64_t wr = batch * std::min<int64_t>(N, p.st_limit < N ? p.st_limit : N);
  ps.bytes = (rd + wr) * esz;
  passes_.push_back(ps);
}

static int64_t prod_cnt(const std::vector<Digit>& d) {
  int64_t p = 1;
  for (auto& x : d) p *= x.cnt;
  return p;
}

// Canonical contiguous layout for the line digits of an intermediate whose
// lines have `len` elements: digit j stride = len * prod(cnt_{>j}).
static std::vector<int64_t> canon_strides(const std::vector<Digit>& d, int64_t len) {
  std::vector<int64_t> s(d.size());
  int64_t acc = len;
  for (int j = (int)d.size() - 1; j >= 0; --j) {
    s[j] = acc;
    acc *= d[j].cnt;
  }
  return s;
}

// Strides of `lines` in an intermediate buffer holding `len` elements per
// line: lines outermost in their order, except that a digit of unit stride on
// both sides (the contiguous rows of a strided N-D axis) stays innermost, so
// the passes on either side keep coalesced line-fast access on it.  *U is
// the intermediate's element stride (that digit's count, or 1).
// *ins is where a new digit of the transform dimension goes in the line
// list (just before that digit, so it can merge with it).
static std::vector<int64_t> inter_strides(const std::vector<Digit>& lines, int64_t len,
                                          int64_t* U, int* ins) {
  int u = -1;
  for (int j = 0; j < (int)lines.size(); ++j)
    if (lines[j].is == 1 && lines[j].os == 1 && lines[j].cnt > 1) u = j;
  *U = u >= 0 ? lines[u].cnt : 1;
  *ins = u >= 0 ? u : (int)lines.size();
  std::vector<Digit> other;
  std::vector<int> idx;
  for (int j = 0; j < (int)lines.size(); ++j)
    if (j != u) {
      other.push_back(lines[j]);
      idx.push_back(j);
    }
  const auto co = canon_strides(other, len * *U);
  std::vector<int64_t> cs(lines.size(), 1);
  for (size_t k = 0; k < idx.size(); ++k) cs[idx[k]] = co[k];
  return cs;
}

static int64_t choose_n1(int64_t N, int64_t cap) {
  // Number of levels so that every factor fits; then the largest smooth
  // divisor not above the balanced size N^(1/levels).
  int levels = 2;
  while (true) {
    long double f = powl((long double)N, 1.0L / levels);
    if (f <= (long double)cap * 1.0000001L) break;
    ++levels;
  }
  const long double target = powl((long double)N, 1.0L / levels);
  int64_t best = 0;
  for (int64_t d = 2; d <= cap && d <= N; ++d) {
    if (N % d || !is_smooth7(d)) continue;
    if ((long double)d <= target * 1.0000001L) best = d;
  }
  if (best == 0) {
    for (int64_t d = std::min(cap, N); d >= 2; --d)
      if (N % d == 0 && is_smooth7(d)) {
        best = d;
        break;
      }
  }
  return best;
}

void PassProgram::emit(int64_t N, std::vector<Digit> lines, int64_t iks, int64_t oks,
                       int sign, const ElemHook& ld, const ElemHook& st) {
  const int64_t cap = lds_cap(prec_);
  const int64_t lfcap = lf_split_ ? lf_cap(prec_) : cap;
  const int64_t L = prod_cnt(lines);
  const bool strided = iks != 1 || oks != 1;
  if ((strided ? (lf_split_ ? strided_single(N, prec_) : N <= cap) : N <= cap) && is_smooth7(N)) {
    emit_single(N, lines, iks, oks, sign, ld, st, nullptr, -1, 0, L * N);
    return;
  }
  if (!is_smooth7(N)) fail(ERR_INTERNAL, "emit: non-smooth length reached four-step");
  // Four-step N = N1 * N2; both passes have a strided side.
  const int64_t N1 = choose_n1(N, lfcap);
  const int64_t N2 = N / N1;
  if (N1 < 2 || N2 < 2) fail(ERR_NOT_IMPL, "no four-step factorisation for this length");

  // Intermediate layout [lines][k1][n2][u] (u: a unit-stride row digit, or
  // none); U is the stride of n2 in it.
  int64_t U = 1;
  int ins = 0;
  const std::vector<int64_t> cs = inter_strides(lines, N, &U, &ins);
  auto shift = [ins](int lv) { return lv >= ins ? lv + 1 : lv; };
  // Pass A: N1-point DFTs along n1 (element stride N2*iks), one line per
  // (lines..., n2); output k1 at stride N2*U with W_N^{n2 k1}.
  std::vector<Digit> la = lines;
  for (size_t j = 0; j < la.size(); ++j) la[j].os = cs[j];
  la.insert(la.begin() + ins, Digit{N2, iks, U});
  const int n2_level = ins;
  // Parent index m = n * kmul + (digit terms); in pass A, n = n1 * N2 + n2.
  ElemHook lda;
  if (ld.active()) {
    if (ld.level2 >= 0) fail(ERR_NOT_IMPL, "load hook deeper than two split levels");
    lda = ld;
    if (lda.level >= 0) lda.level = shift(lda.level);
    lda.kmul = N2 * ld.kmul;
    if (ld.level >= 0) {
      lda.level2 = n2_level;
      lda.dmul2 = ld.kmul;
    } else {
      lda.level = n2_level;
      lda.dmul = ld.kmul;
    }
  }
  emit_single(N1, la, N2 * iks, N2 * U, sign, lda, ElemHook{}, nullptr, n2_level, N, L * N);

  // Pass B: N2-point DFTs along n2 (stride U in the intermediate), one line
  // per (lines..., k1); the store writes X[k1 + N1 k2] in the caller layout.
  std::vector<Digit> lb = lines;
  for (size_t j = 0; j < lb.size(); ++j) lb[j].is = cs[j];
  lb.insert(lb.begin() + ins, Digit{N1, N2 * U, oks});
  // Parent index m = k * kmul + (digit terms) with k = k1 + N1 * k2.
  ElemHook stb;
  if (st.active()) {
    if (st.level2 >= 0) fail(ERR_NOT_IMPL, "store hook deeper than two split levels");
    stb = st;
    stb.kmul = N1 * st.kmul;
    if (st.level >= 0) {
      stb.level = shift(st.level);
      stb.level2 = ins;
      stb.dmul2 = st.kmul;
    } else {
      stb.level = ins;
      stb.dmul = st.kmul;
    }
  }
  emit(N2, lb, U, N1 * oks, sign, ElemHook{}, stb);
}

void PassProgram::add_transform(int64_t N, const std::vector<Digit>& lines, int64_t iks,
                                int64_t oks) {
  const size_t p0 = passes_.size();
  lf_split_ = true;
  try {
    add_transform_impl(N, lines, iks, oks);
  } catch (const Error& e) {
    if (e.status != ERR_NOT_IMPL) throw;
    // Splitting strided lines at the coalescing cap can need more line
    // digits than a pass has (deep Bluestein chains on strided axes): plan
    // that transform with the LDS cap alone.  Tables already uploaded stay.
    passes_.resize(p0);
    lf_split_ = false;
    add_transform_impl(N, lines, iks, oks);
  }
  lf_split_ = true;
}

void PassProgram::add_transform_impl(int64_t N, const std::vector<Digit>& lines, int64_t iks,
                                     int64_t oks) {
  if (N < 1) fail(ERR_INVALID_DIM, "transform length < 1");
  if (N == 1) return;  // identity
  const int64_t cap =
      (lf_split_ && (iks != 1 || oks != 1)) ? lf_cap(prec_) : lds_cap(prec_);
  const int64_t L = prod_cnt(lines);
  if (is_smooth7(N)) {
    emit(N, lines, iks, oks, 0, ElemHook{}, ElemHook{});
    return;
  }
  const int64_t M = next_pow2(2 * N - 1);
  // Direction-dependent chirp tables: [0] forward, [1] backward.
  std::vector<cld> chirp[2], H[2];
  bluestein_tables_pair(N, M, chirp, H);
  const void* tc[2] = {upload(chirp[0]), upload(chirp[1])};
  const size_t first = passes_.size();
  if (M <= cap) {
    const void* th[2] = {upload(H[0]), upload(H[1])};
    ElemHook ld, st;
    ld.table = tc[0];
    ld.limit = N;
    st.table = tc[0];
    st.limit = N;
    emit_single(M, lines, iks, oks, -1, ld, st, th[0], -1, 0, L * N);
    PassSpec& ps = passes_.back();
    ps.sign = 0;
    ps.p.n = (int32_t)M;
    ps.dir_tables = true;
    ps.pre_t[0] = tc[0];
    ps.pre_t[1] = tc[1];
    ps.post_t[0] = tc[0];
    ps.post_t[1] = tc[1];
    ps.blu_t[0] = th[0];
    ps.blu_t[1] = th[1];
    return;
  }
  if (iks == 1 && oks == 1 && env_int("SHAFFT_AMD_BLUE3", 1) &&
      bluestein_three_pass(N, M, lines, tc, H)) {
    return;
  }
  // The filter in natural order (the three-pass form uploads its own
  // transposed copy).
  const void* th[2] = {upload(H[0]), upload(H[1])};
  // Long Bluestein: forward M-FFT (chirp + zero padding on its first load,
  // filter product on its last store), then backward M-FFT (output chirp and
  // truncation on its last store).  The intermediate is [lines][M][u] as
  // in the four-step (u: a unit-stride row digit, or none).
  std::vector<Digit> lf = lines;
  int64_t U = 1;
  int ins = 0;
  const std::vector<int64_t> cs = inter_strides(lines, M, &U, &ins);
  for (size_t j = 0; j < lf.size(); ++j) lf[j].os = cs[j];
  ElemHook ld;
  ld.table = tc[0];
  ld.limit = N;
  ElemHook sth;
  sth.table = th[0];
  emit(M, lf, iks, U, -1, ld, sth);
  std::vector<Digit> lbk = lines;
  for (size_t j = 0; j < lbk.size(); ++j) lbk[j].is = cs[j];
  ElemHook st;
  st.table = tc[0];
  st.limit = N;
  emit(M, lbk, U, oks, +1, ElemHook{}, st);
  // Mark direction-dependent tables and intermediate sizes.
  for (size_t i = first; i < passes_.size(); ++i) {
    PassSpec& ps = passes_[i];
    ps.out_elems = (i + 1 == passes_.size()) ? L * N : L * M;
    if (ps.p.pre == tc[0] || ps.p.post == tc[0] || ps.p.post == th[0]) {
      ps.dir_tables = true;
      ps.pre_t[0] = ps.p.pre;
      ps.pre_t[1] = ps.p.pre == tc[0] ? tc[1] : ps.p.pre;
      ps.post_t[0] = ps.p.post;
      ps.post_t[1] = ps.p.post == tc[0] ? tc[1] : (ps.p.post == th[0] ? th[1] : ps.p.post);
      ps.blu_t[0] = ps.blu_t[1] = nullptr;
    }
  }
}

// Move a pass of the three-pass Bluestein to the compile-time-length kernels
// of fft_blue.hip: the kernel's radix plan and stage table, C = 2^logC lines
// per tile, one tile per workgroup.  Column passes (line-fast on both sides)
// take 256-byte row segments (32 complex-float / 16 complex-double adjacent
// columns); row passes take at least 256 threads.  Knobs:
//   SHAFFT_AMD_BLUE_FX=0   keep the generic kernel
//   SHAFFT_AMD_BLUE_COLC   columns per column tile
//   SHAFFT_AMD_BLUE_ROWT   minimum threads per row tile
//   SHAFFT_AMD_BLUE_TWL    stage twiddles staged in LDS: 1 always, 0 never
//                          (default: column passes only; a row pass's table
//                          of ~n entries would cost workgroups per CU:
//                          1,000,003-point complex double row pass 37.2 us
//                          with it, 26.0 us without)
bool PassProgram::configure_blue(PassSpec& ps, int kind) {
  sa_pass_t& p = ps.p;
  const int n = p.n;
  if (!env_int("SHAFFT_AMD_BLUE_FX", 1) || !sa_blue_supported(n, kind)) return false;
  const bool cols = kind != SA_BLUE_ROWS;
  if (cols ? (p.ld_kfast || p.st_kfast) : (!p.ld_kfast || !p.st_kfast)) return false;
  const int esz = prec_ ? 16 : 8;
  // Column tiles: 256-byte row segments while that leaves a tile for every
  // CU, else narrower (not below 64 bytes): these passes are short, and a
  // CU without a tile costs more than the partial cache lines of narrower
  // segments (1,000,003-point complex float, 4096 columns: 16 columns,
  // 13.1 us, against 32 columns on half the CUs, 15.5 us).
  int C = 1;
  if (cols) {
    C = 256 / esz;
    while (C > 64 / esz && (p.batch + C - 1) / C < device_cus()) C /= 2;
    C = env_int("SHAFFT_AMD_BLUE_COLC", C);
  } else {
    C = std::max(1, env_int("SHAFFT_AMD_BLUE_ROWT", 256) / (n / 16));
  }
  // Complex-float row passes with fewer rows than CUs take 8 elements per
  // thread (twice the threads per row, one more stage): 16 / 32 / 64 rows
  // of 4096 points 8.9 -> 8.0, 9.0 -> 8.2, 9.2 -> 8.2 us; with a row for
  // every CU the extra stage cancels the gain (1,000,003 points, 512 rows:
  // 15.5 / 15.8 us), and complex-double rows gain nothing (40,009 points:
  // 11.1 / 11.3 us).  SHAFFT_AMD_BLUE_ROW_E=16/8 forces either form.
  int E = sa_blue_elems(n, prec_, kind, C);
  if (!cols && n <= 4096) {
    const int knob = env_int("SHAFFT_AMD_BLUE_ROW_E", 0);
    if (knob == 8 || (knob == 0 && prec_ == 0 && p.batch < device_cus())) E = 8;
  }
  const int T = n / E;
  int logC = 0;
  while ((1 << (logC + 1)) <= C) ++logC;
  while (logC > 0 && ((int64_t)1 << (logC - 1)) >= p.batch) --logC;
  int stride = n + (n - 1) / 16;
  if (stride % 2 == 0) stride += 1;
  const int budget = lds_budget_bytes();
  while (logC > 0 && (((int64_t)T << logC) > 512 || ((int64_t)stride << logC) * esz > budget)) --logC;
  if (((int64_t)T << logC) > 512 || ((int64_t)stride << logC) * esz > budget) return false;
  const std::vector<int> rad = fixed_radix_plan(n, E);
  if (rad.empty()) return false;
  std::vector<int> offs;
  p.tw = stage_table(rad, offs);
  p.nstages = (int32_t)rad.size();
  int64_t Ns = 1;
  for (size_t s = 0; s < rad.size(); ++s) {
    p.radix[s] = rad[s];
    p.ns[s] = (int32_t)Ns;
    p.twoff[s] = offs[s];
    p.nsdiv[s] = make_fastdiv((uint32_t)Ns);
    p.nbdiv[s] = make_fastdiv((uint32_t)(n / rad[s]));
    Ns *= rad[s];
  }
  int64_t tw_elems = 0;
  for (int s = 1; s < p.nstages; ++s)
    tw_elems = std::max<int64_t>(tw_elems, (int64_t)p.twoff[s] + (int64_t)p.ns[s] * (p.radix[s] - 1));
  p.tw_lds = 0;
  if (env_int("SHAFFT_AMD_BLUE_TWL", cols ? 1 : 0) && tw_elems > 0 &&
      (((int64_t)stride << logC) + tw_elems) * esz <= budget)
    p.tw_lds = (int32_t)tw_elems;
  p.logC = logC;
  p.C = 1 << logC;
  p.T = T;
  p.emax = E;
  p.ldsstride = stride;
  p.Tdiv = make_fastdiv((uint32_t)T);
  p.Cdiv = make_fastdiv((uint32_t)p.C);
  p.reg_first = p.reg_last = 0;
  p.xcd_order = 0;
  p.lds_pad = 1;  // the Bluestein kernels' rows are padded
  ps.blue = kind;
  return true;
}

// Bluestein convolution of length M = N1 * N2 in three passes over a
// contiguous axis (iks = oks = 1), keeping the spectrum in transposed order:
//   A: N1-point forward DFTs along n1 (element stride N2) of the chirped,
//      zero-padded input, times W_M^{n2 k1}, stored as Y[k1][n2];
//   B: per row k1, the N2-point forward DFT over n2, times the filter
//      H[k1 + N1 k2] (a transposed copy, one row per k1), the N2-point
//      inverse DFT over k2, times conj(W_M^{n2 k1}) -- all in LDS;
//   C: N1-point inverse DFTs along k1, times the output chirp, stored for
//      n = N2 n1 + n2 < N.
// The inverse four-step x[N2 n1 + n2] = sum_k1 w_N1^{-n1 k1} w_M^{-n2 k1}
// sum_k2 X[k1 + N1 k2] w_N2^{-n2 k2} needs X only row by row, so the
// forward transform's final transpose and the inverse's first one cancel:
// 3 passes over M instead of the 6 of two separate M-point transforms.
bool PassProgram::bluestein_three_pass(int64_t N, int64_t M, const std::vector<Digit>& lines,
                                       const void* const tc[2],
                                       const std::vector<std::complex<long double>> H[2]) {
  // Column passes of at most 512 points (16 adjacent columns of the
  // chirped input per tile at the least).
  const int64_t lfc = std::min<int64_t>(lf_cap(prec_), 512);
  int64_t N2 = std::min<int64_t>(env_int("SHAFFT_AMD_BLUE_N2", 4096), M / 16);
  int64_t N1 = M / N2;
  if (N1 > lfc) {
    N1 = lfc;
    N2 = M / N1;
  }
  if (N1 < 2 || N2 < 2 || N1 * N2 != M || N1 > lfc || N2 > std::min<int64_t>(8192, lds_cap(prec_)))
    return false;
  int64_t U = 1;
  int ins = 0;
  const std::vector<int64_t> cs = inter_strides(lines, M, &U, &ins);
  if (U != 1) return false;
  const int64_t L = prod_cnt(lines);
  std::vector<cld> Ht[2];
  for (int d = 0; d < 2; ++d) {
    Ht[d].resize(M);
    for (int64_t k1 = 0; k1 < N1; ++k1)
      for (int64_t k2 = 0; k2 < N2; ++k2) Ht[d][k1 * N2 + k2] = H[d][k1 + N1 * k2];
  }
  const void* th[2] = {upload(Ht[0]), upload(Ht[1])};
  const size_t first = passes_.size();

  std::vector<Digit> la = lines;
  for (size_t j = 0; j < la.size(); ++j) la[j].os = cs[j];
  la.push_back(Digit{N2, 1, 1});
  const int n2a = (int)la.size() - 1;
  ElemHook lda;
  lda.table = tc[0];
  lda.limit = N;
  lda.kmul = N2;
  lda.level = n2a;
  lda.dmul = 1;
  emit_single(N1, la, N2, N2, -1, lda, ElemHook{}, nullptr, n2a, M, L * M);
  configure_blue(passes_.back(), SA_BLUE_COLS_IN);

  std::vector<Digit> lb = lines;
  for (size_t j = 0; j < lb.size(); ++j) lb[j].is = lb[j].os = cs[j];
  lb.push_back(Digit{N1, N2, N2});
  const int k1b = (int)lb.size() - 1;
  emit_single(N2, lb, 1, 1, -1, ElemHook{}, ElemHook{}, th[0], k1b, M, L * M);
  {
    sa_pass_t& p = passes_.back().p;
    p.ptw_conj = 1;
    p.blu_level = p.ptw_level;  // the k1 digit, after emit_single's remapping
    p.blu_hstride = N2;
  }
  configure_blue(passes_.back(), SA_BLUE_ROWS);

  std::vector<Digit> lc = lines;
  for (size_t j = 0; j < lc.size(); ++j) lc[j].is = cs[j];
  lc.push_back(Digit{N2, 1, 1});
  const int n2c = (int)lc.size() - 1;
  ElemHook stc;
  stc.table = tc[0];
  stc.limit = N;
  stc.kmul = N2;
  stc.level = n2c;
  stc.dmul = 1;
  emit_single(N1, lc, N2, N2, +1, ElemHook{}, stc, nullptr, -1, 0, L * N);
  configure_blue(passes_.back(), SA_BLUE_COLS_OUT);

  PassSpec* ps = &passes_[first];
  for (int d = 0; d < 3; ++d) ps[d].dir_tables = true;
  ps[0].pre_t[0] = tc[0];
  ps[0].pre_t[1] = tc[1];
  ps[1].blu_t[0] = th[0];
  ps[1].blu_t[1] = th[1];
  ps[2].post_t[0] = tc[0];
  ps[2].post_t[1] = tc[1];
  return true;
}

PassProgram::SlotMap PassProgram::assign(Mode mode, int64_t data_elems) const {
  SlotMap m;
  const size_t np = passes_.size();
  m.in.assign(np, -1);
  m.out.assign(np, -1);
  if (np == 0) {
    m.result = SLOT_DATA;
    return m;
  }
  auto use_scratch = [&](int slot, int64_t elems) {
    int i = slot - SLOT_S1;
    m.scratch[i] = std::max(m.scratch[i], elems);
  };
  if (mode == PINGPONG) {
    int cur = SLOT_DATA;
    for (size_t i = 0; i < np; ++i) {
      m.in[i] = cur;
      int out;
      if (passes_[i].out_elems <= data_elems) {
        out = cur == SLOT_DATA ? SLOT_WORK : SLOT_DATA;
      } else {
        out = cur == SLOT_S1 ? SLOT_S2 : SLOT_S1;
        use_scratch(out, passes_[i].out_elems);
      }
      m.out[i] = out;
      cur = out;
    }
    m.result = cur;
    return m;
  }
  // PRESERVE / INPLACE: last pass writes WORK; walk backwards alternating.
  m.out[np - 1] = SLOT_WORK;
  for (int i = (int)np - 2; i >= 0; --i) {
    const int next_out = m.out[i + 1];
    int out;
    if (mode == PRESERVE && passes_[i].out_elems <= data_elems && next_out != SLOT_WORK)
      out = SLOT_WORK;
    else
      out = next_out == SLOT_S1 ? SLOT_S2 : SLOT_S1;
    if (out != SLOT_WORK) use_scratch(out, passes_[i].out_elems);
    m.out[i] = out;
  }
  m.in[0] = SLOT_DATA;
  for (size_t i = 1; i < np; ++i) m.in[i] = m.out[i - 1];
  m.result = SLOT_WORK;
  return m;
}

void PassProgram::run_one(const SlotMap& m, size_t i, void* const slots[SLOT_COUNT], int dir,
                          hipStream_t stream) const {
  const PassSpec& ps = passes_[i];
  sa_pass_t p = ps.p;
  p.in = slots[m.in[i]];
  p.out = slots[m.out[i]];
  if (!p.in || !p.out) fail(ERR_NO_BUFFER, "pass buffer not set");
  if (ps.dir_tables) {
    const int d = dir < 0 ? 0 : 1;
    p.pre = ps.pre_t[d];
    p.post = ps.post_t[d];
    p.blu_h = ps.blu_t[d];
  }
  const int sign = ps.sign == 0 ? dir : ps.sign;
  int rc = ps.blue    ? sa_launch_blue(&p, prec_, ps.blue, stream)
           : ps.fixed ? sa_launch_fixed(&p, prec_, sign, ps.grid, stream)
                      : sa_launch_pass(&p, prec_, sign, stream);
  hip_check((hipError_t)rc, "fft pass launch");
}

void PassProgram::run(const SlotMap& m, void* const slots[SLOT_COUNT], int dir,
                      hipStream_t stream) const {
  for (size_t i = 0; i < passes_.size(); ++i) run_one(m, i, slots, dir, stream);
}

int64_t PassProgram::table_bytes() const {
  int64_t b = 0;
  for (auto& t : tables_) b += (int64_t)t.bytes;
  return b;
}

int64_t PassProgram::total_bytes() const {
  int64_t b = 0;
  for (auto& p : passes_) b += p.bytes;
  return b;
}

std::string PassProgram::describe() const {
  std::ostringstream os;
  for (size_t i = 0; i < passes_.size(); ++i) {
    const sa_pass_t& p = passes_[i].p;
    os << "pass " << i << ": n=" << p.n << " batch=" << p.batch
       << (passes_[i].fixed ? " fixed(grid=" + std::to_string(passes_[i].grid) + ")" : "")
       << (passes_[i].blue ? " blue" + std::to_string(passes_[i].blue) : "")
       << " radices=";
    for (int s = 0; s < p.nstages; ++s) os << (s ? "x" : "") << p.radix[s];
    os << " iks=" << p.iks << " oks=" << p.oks << " C=" << p.C << " T=" << p.T
       << " ld=" << (p.ld_kfast ? "k" : "line") << " st=" << (p.st_kfast ? "k" : "line")
       << (p.blu_h || passes_[i].blu_t[0] ? " bluestein" : "")
       << (p.ptw_level >= 0 ? " twiddle" : "") << (p.lds_pad ? "" : " unpadded");
    if (p.ld_level >= 0 || p.ld_limit < p.n || p.pre || passes_[i].pre_t[0])
      os << " ld_hook(kmul=" << p.ld_kmul << ",lvl=" << p.ld_level << "/" << p.ld_level2
         << ",lim=" << p.ld_limit << ")";
    if (p.st_level >= 0 || p.st_limit < p.n || p.post || passes_[i].post_t[0])
      os << " st_hook(kmul=" << p.st_kmul << ",lvl=" << p.st_level << "/" << p.st_level2
         << ",lim=" << p.st_limit << ")";
    os << "\n";
  }
  return os.str();
}

}  // namespace sa
