// Pass planner: turns "a length-N DFT over a set of strided lines" into a
// short list of LDS-resident passes (pass_desc.h) and runs them.
//
//   * N fits in LDS and is 7-smooth      -> one Stockham pass.
//   * N fits but has a larger prime      -> one in-LDS Bluestein pass.
//   * N too long, 7-smooth               -> four-step N = N1 * N2: pass A does
//     the N1-point DFTs with the W_N^{n2 k1} twiddle fused into its store,
//     then the N2-point sub-problem runs with the (k1, k2) -> k1 + N1 k2
//     transpose fused into its final store.  Recursive.
//   * N too long, not smooth             -> Bluestein over a length M = N1*N2
//     (power of two) in three passes: strided N1-point columns with the chirp
//     and zero padding fused into the load; the N2-point rows doing forward
//     DFT, filter product and inverse DFT in LDS (spectrum left transposed);
//     inverse N1-point columns with the output chirp and truncation fused into
//     the store.  Falls back to two separate M-point four-step chains.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <complex>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../kernels/pass_desc.h"
#include "common.hpp"

namespace sa {

// Owned device allocation.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  explicit DevBuf(size_t n);
  ~DevBuf();
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) {
    o.p = nullptr;
    o.bytes = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept;
};

// One line-digit of a pass: count, stride on the input side, stride on the
// output side (complex elements).
struct Digit {
  int64_t cnt, is, os;
};

// Buffer slots a pass reads/writes.
enum Slot : int { SLOT_DATA = 0, SLOT_WORK = 1, SLOT_S1 = 2, SLOT_S2 = 3, SLOT_COUNT = 4 };

struct PassSpec {
  sa_pass_t p{};
  int sign = 0;  // 0: direction of the execute call, else fixed -1/+1
  int64_t out_elems = 0;  // elements of the buffer it writes (for slot sizing)
  int64_t bytes = 0;  // algorithmic HBM bytes (one read + one write per element)
  // Fixed-length persistent kernel and its grid (workgroups).
  bool fixed = false;
  int grid = 0;
  // Three-pass Bluestein kernel kind (SA_BLUE_*), 0: none.
  int blue = 0;
  // Bluestein passes carry direction-dependent tables: [0] forward, [1] backward.
  bool dir_tables = false;
  const void* pre_t[2] = {nullptr, nullptr};
  const void* post_t[2] = {nullptr, nullptr};
  const void* blu_t[2] = {nullptr, nullptr};
};

// Per-element index hooks used by the Bluestein chains.
struct ElemHook {
  const void* table = nullptr;  // pre/post factor table (device)
  int64_t limit = -1;           // valid iff index < limit (-1: all)
  int64_t kmul = 1;  // index = k * kmul + digit[level] * dmul + digit[level2] * dmul2
  int level = -1;    // index into the line digits, -1: none
  int64_t dmul = 0;
  int level2 = -1;
  int64_t dmul2 = 0;
  bool active() const { return table != nullptr || limit >= 0; }
};

// Precision-dependent sizes.
struct Prec {
  int prec;  // 0 complex float, 1 complex double
  int bytes() const { return prec ? 16 : 8; }
};

// Largest LDS-resident line for a precision (elements).
int64_t lds_cap(int prec);
// Longest line for a pass with strided elements (coalescing bound).
int64_t lf_cap(int prec);
bool is_smooth7(int64_t n);
// radices for a 7-smooth n (product == n), none above maxrad (the generic
// kernel has radices up to 21 for complex float, 28 for complex double)
std::vector<int> radix_plan(int64_t n, int maxrad = 28);

class PassProgram {
 public:
  explicit PassProgram(int prec) : prec_(prec) {}
  // Append the passes of a length-N DFT over the lines `lines`
  // (outermost first), input element stride iks, output element stride oks.
  void add_transform(int64_t N, const std::vector<Digit>& lines, int64_t iks, int64_t oks);
  // Buffer-slot assignment for one execution mode:
  //   PINGPONG  alternate DATA/WORK (input may be overwritten), result in either;
  //   PRESERVE  never write DATA, result in WORK;
  //   INPLACE   DATA and WORK alias: only the last pass writes WORK.
  // Intermediates longer than `data_elems` always go to scratch S1/S2.
  enum Mode { PINGPONG = 0, PRESERVE = 1, INPLACE = 2 };
  struct SlotMap {
    std::vector<int> in, out;
    int result = SLOT_DATA;
    int64_t scratch[2] = {0, 0};
  };
  SlotMap assign(Mode mode, int64_t data_elems) const;
  // Launch every pass.  slots[i] are device pointers; dir -1/+1.
  void run(const SlotMap& m, void* const slots[SLOT_COUNT], int dir, hipStream_t stream) const;
  // Launch a single pass (profiling aid).
  void run_one(const SlotMap& m, size_t i, void* const slots[SLOT_COUNT], int dir,
               hipStream_t stream) const;

  size_t npasses() const { return passes_.size(); }
  PassSpec& pass_mut(size_t i) { return passes_[i]; }
  const PassSpec& pass(size_t i) const { return passes_[i]; }
  int64_t total_bytes() const;
  int64_t table_bytes() const;  // device bytes of the plan's tables
  // Elements pass i reads / writes when the program's input has data_elems.
  int64_t pass_in_elems(size_t i, int64_t data_elems) const {
    return i == 0 ? data_elems : passes_[i - 1].out_elems;
  }
  int64_t pass_out_elems(size_t i) const { return passes_[i].out_elems; }
  std::string describe() const;
  int precision() const { return prec_; }

 private:
  int prec_;
  std::vector<PassSpec> passes_;
  void add_transform_impl(int64_t N, const std::vector<Digit>& lines, int64_t iks, int64_t oks);
  bool bluestein_three_pass(int64_t N, int64_t M, const std::vector<Digit>& lines,
                            const void* const tc[2],
                            const std::vector<std::complex<long double>> H[2]);
  bool configure_blue(PassSpec& ps, int kind);
  bool lf_split_ = true;  // split strided lines at lf_cap (else only at lds_cap)
  std::vector<DevBuf> tables_;
  std::map<std::vector<int>, std::pair<const void*, std::vector<int>>> stage_cache_;

  const void* upload(const std::vector<std::complex<long double>>& t);
  const void* stage_table(const std::vector<int>& radices, std::vector<int>& offsets);

  void emit(int64_t N, std::vector<Digit> lines, int64_t iks, int64_t oks, int sign,
            const ElemHook& ld, const ElemHook& st);
  void emit_single(int64_t N, std::vector<Digit> lines, int64_t iks, int64_t oks, int sign,
                   const ElemHook& ld, const ElemHook& st, const void* blu_h, int ptw_level,
                   int64_t ptw_N, int64_t out_elems);
};

}  // namespace sa
