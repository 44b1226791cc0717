// Backend objects behind the seam the reference core calls
// (src/cpp/fft/hipfft_method/fftnd_method.hpp:14-31,
//  src/cpp/fft/hipfft_method/fft1d_method.hpp:14-43,
//  src/cpp/fft/hipfft_method/common/normalize.hpp).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <vector>

#include "planner.hpp"

namespace sa {

// N-D local transform over a consecutive block of axes of a row-major
// tensor (the reference's one-rank FFTND subplan, fftnd_method.cpp:162-351).
class NDBackend {
 public:
  NDBackend(int nta, const int* ta, int ndim, const int* dims, int prec);
  // Out-of-place data -> work; afterwards `data` holds the result (pointers
  // swapped when the last pass wrote `work`).  Host-synchronous.
  void execute(void*& data, void*& work, int dir);
  void set_stream(hipStream_t s) { stream_ = s; }
  hipStream_t stream() const { return stream_; }
  int prec() const { return prec_; }
  int64_t elems() const { return elems_; }
  const PassProgram& program() const { return prog_; }
  // Profiling aid: run pass i alone from `in` to `out` (asynchronous); the
  // buffers' element counts must cover the pass's extents.
  void execute_pass(size_t i, void* in, int64_t in_elems, void* out, int64_t out_elems, int dir);
  void pass_extents(size_t i, int64_t* in_elems, int64_t* out_elems) const;
  // in -> out with `in` preserved, asynchronous on `stream`, using the
  // caller's scratch s1/s2 (preserve_scratch_elems(0/1) elements).  The plan
  // holds no per-call state, so calls on different streams may overlap.
  void execute_preserve(const void* in, void* out, void* s1, void* s2, int dir,
                        hipStream_t stream) const;
  int64_t preserve_scratch_elems(int i) const { return map_pr_.scratch[i]; }
  const PassProgram::SlotMap& preserve_map() const { return map_pr_; }
  int64_t table_bytes() const { return prog_.table_bytes(); }

 private:
  static int rotate_mode();  // SHAFFT_AMD_ROTATE: 0 never, 1 always, -1 automatic
  void plan_rotated(const std::vector<int>& blk, int ndim, const int* dims);
  int prec_;
  int64_t elems_ = 1;
  PassProgram prog_;
  PassProgram::SlotMap map_pp_, map_ip_, map_pr_;
  DevBuf s1_, s2_;
  hipStream_t stream_ = nullptr;
};

// 1-D transform of one local line at one rank (the reference's FFT1D
// backend; fft1d_method.hip:416-650).  Input is preserved.
class OneDBackend {
 public:
  // own_scratch=false leaves scratch to the caller (execute_preserve only).
  OneDBackend(int64_t n, int prec, bool own_scratch = true);
  void execute(void* in, void* out, int dir);
  void execute_preserve(const void* in, void* out, void* s1, void* s2, int dir,
                        hipStream_t stream) const;
  // execute() with the intermediates in the caller's `scratch` of
  // `scratch_bytes` (no allocation); fails with ERR_ALLOC if the plan needs
  // more.  Asynchronous on the plan's stream.
  void execute_in_scratch(const void* in, void* out, int dir, void* scratch,
                          size_t scratch_bytes) const;
  int64_t preserve_scratch_elems(int i) const { return map_pr_.scratch[i]; }
  const PassProgram::SlotMap& preserve_map() const { return map_pr_; }
  int64_t table_bytes() const { return prog_.table_bytes(); }
  void set_stream(hipStream_t s) { stream_ = s; }
  hipStream_t stream() const { return stream_; }
  int prec() const { return prec_; }
  int64_t n() const { return n_; }
  const PassProgram& program() const { return prog_; }
  void execute_pass(size_t i, void* in, int64_t in_elems, void* out, int64_t out_elems, int dir);
  void pass_extents(size_t i, int64_t* in_elems, int64_t* out_elems) const;
  int64_t scratch_bytes() const { return (int64_t)(s1_.bytes + s2_.bytes); }

 private:
  int64_t n_;
  int prec_;
  PassProgram prog_;
  PassProgram::SlotMap map_pr_, map_ip_;
  DevBuf s1_, s2_;
  hipStream_t stream_ = nullptr;
};

// x *= f over `count` complex elements on `stream` (asynchronous).
void scale_async(void* data, double f, uint64_t count, int prec, hipStream_t stream);

}  // namespace sa
