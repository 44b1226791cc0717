#include "backend.hpp"

#include <algorithm>
#include <cstdlib>
#include <sstream>
#include <string>

namespace sa {

// Grow s1/s2 to what slot map `m` needs (allocation happens at plan time
// for the primary mode and on first use for the aliased-buffer mode).
static void ensure_scratch(const PassProgram::SlotMap& m, int prec, DevBuf& s1, DevBuf& s2) {
  const size_t esz = prec ? 16 : 8;
  if ((size_t)m.scratch[0] * esz > s1.bytes) s1 = DevBuf((size_t)m.scratch[0] * esz);
  if ((size_t)m.scratch[1] * esz > s2.bytes) s2 = DevBuf((size_t)m.scratch[1] * esz);
}

// SHAFFT_AMD_ROTATE: 0 never, 1 always, unset: when the axis-by-axis plan's
// outermost pass would read and write complex64 rows 4 MiB or more apart.
// Its tiles are 16 rows wide there (128-byte segments; 32 rows of 1024
// points do not fit LDS), and on MI355X such a pass runs at ~3 TB/s (1024^3:
// 3.06 TB/s; a plain copy with the same tiles 3.3 TB/s) while the rotated
// plan's strided side is one axis length; measured 1024^3 complex64:
// 12.04 -> 10.72 ms.  Complex double tiles are 256 bytes wide and the
// axis-by-axis plan is faster there (512^3: 2.52 vs 2.71 ms).
int NDBackend::rotate_mode() {
  const char* e = std::getenv("SHAFFT_AMD_ROTATE");
  return (e && *e) ? std::atoi(e) : -1;
}

// Rotated plan for a block of transformed axes that ends at the last (unit
// stride) axis.  The block's layout is rotated left by one axis per pass:
// with block axes a1..am and layout L0 = [a1..am], pass t transforms the
// innermost axis of L(t-1) and writes L(t) = rotl(L(t-1)), in which that
// axis is second innermost and the next axis to transform is innermost.
// After m passes the layout is L0 again.  Every pass therefore reads its
// lines contiguously and writes them with the element stride of one axis
// length, lines adjacent (no pass sees the block's largest stride on both
// sides, as the axis-by-axis plan's last pass does).  Leading untouched
// axes stay outermost.
void NDBackend::plan_rotated(const std::vector<int>& blk, int ndim, const int* dims) {
  const int m = (int)blk.size();
  int64_t M = 1;
  for (int ax : blk) M *= dims[ax];
  const int64_t outer = elems_ / M;
  std::vector<int> cur(m);
  for (int i = 0; i < m; ++i) cur[i] = i;  // layout: block axis indices, outermost first
  auto strides = [&](const std::vector<int>& lay) {
    std::vector<int64_t> st(m);
    int64_t acc = 1;
    for (int i = m - 1; i >= 0; --i) {
      st[lay[i]] = acc;
      acc *= dims[blk[lay[i]]];
    }
    return st;
  };
  for (int t = 0; t < m; ++t) {
    std::vector<int> nxt(cur.begin() + 1, cur.end());
    nxt.push_back(cur[0]);
    const std::vector<int64_t> si = strides(cur), so = strides(nxt);
    const int x = cur.back();
    std::vector<Digit> lines;
    if (outer > 1) lines.push_back(Digit{outer, M, M});
    for (int i = 0; i + 1 < m; ++i) {
      const int y = cur[i];
      lines.push_back(Digit{dims[blk[y]], si[y], so[y]});
    }
    prog_.add_transform(dims[blk[x]], lines, si[x], so[x]);
    cur = nxt;
  }
  (void)ndim;
}

NDBackend::NDBackend(int nta, const int* ta, int ndim, const int* dims, int prec)
    : prec_(prec), prog_(prec) {
  if (prec != 0 && prec != 1) fail(ERR_INVALID_FFTTYPE, "invalid FFT type");
  if (ndim < 1 || !dims) fail(ERR_INVALID_DIM, "invalid ndim");
  for (int i = 0; i < ndim; ++i)
    if (dims[i] <= 0) fail(ERR_INVALID_DIM, "dimension must be positive");
  if (nta < 1 || nta > ndim || !ta) fail(ERR_INVALID_DIM, "invalid transform axes");
  std::vector<int> axes(ta, ta + nta);
  std::sort(axes.begin(), axes.end());
  for (int i = 0; i < nta; ++i) {
    if (axes[i] < 0 || axes[i] >= ndim) fail(ERR_INVALID_DIM, "transform axis out of range");
    // Same restriction as the reference seam (fftnd_method.cpp:89-97).
    if (i + 1 < nta && axes[i] + 1 != axes[i + 1])
      fail(ERR_INVALID_DIM, "transform axes must be consecutive");
  }
  for (int i = 0; i < ndim; ++i) elems_ *= dims[i];
  // Transformed axes of length > 1 (a length-1 axis moves no data).
  std::vector<int> blk;
  for (int ax : axes)
    if (dims[ax] > 1) blk.push_back(ax);
  const bool last_axis = axes.back() == ndim - 1;
  bool rotate = false;
  if (last_axis && blk.size() >= 2) {
    const int mode = rotate_mode();
    int64_t outer_stride = 1;
    for (int i = blk[0] + 1; i < ndim; ++i) outer_stride *= dims[i];
    rotate = mode == 1 || (mode < 0 && prec == 0 && outer_stride * 8 >= (int64_t(4) << 20));
  }
  if (rotate) {
    plan_rotated(blk, ndim, dims);
  } else {
    // Fastest axis first: the contiguous pass runs on the freshest data.
    for (int j = nta - 1; j >= 0; --j) {
      const int ax = axes[j];
      int64_t outer = 1, inner = 1;
      for (int i = 0; i < ax; ++i) outer *= dims[i];
      for (int i = ax + 1; i < ndim; ++i) inner *= dims[i];
      const int64_t n = dims[ax];
      std::vector<Digit> lines = {{outer, n * inner, n * inner}, {inner, 1, 1}};
      prog_.add_transform(n, lines, inner, inner);
    }
  }
  map_pp_ = prog_.assign(PassProgram::PINGPONG, elems_);
  map_ip_ = prog_.assign(PassProgram::INPLACE, elems_);
  map_pr_ = prog_.assign(PassProgram::PRESERVE, elems_);
  ensure_scratch(map_pp_, prec_, s1_, s2_);
  if (diag_enabled()) {
    std::ostringstream os;
    os << "ndim=" << ndim << " nta=" << nta << " prec=" << (prec ? "Z2Z" : "C2C") << " passes=\n"
       << prog_.describe();
    diag("fftndPlan", os.str());
  }
}

void NDBackend::execute(void*& data, void*& work, int dir) {
  if (!data || !work) fail(ERR_NO_BUFFER, "data/work buffer not set");
  const bool inplace = data == work;
  const PassProgram::SlotMap& m = inplace ? map_ip_ : map_pp_;
  if (inplace) ensure_scratch(m, prec_, s1_, s2_);
  void* slots[SLOT_COUNT] = {data, work, s1_.p, s2_.p};
  prog_.run(m, slots, dir, stream_);
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  if (!inplace && m.result == SLOT_WORK) std::swap(data, work);
}

// Shared by both backends: run a PRESERVE-mode program in -> out.
static void run_preserve(const PassProgram& prog, const PassProgram::SlotMap& m, int64_t elems,
                         int prec, const void* in, void* out, void* s1, void* s2, int dir,
                         hipStream_t stream) {
  if (!in || !out) fail(ERR_NO_BUFFER, "buffer not set");
  if (in == out) fail(ERR_INVALID_LAYOUT, "input-preserving execute needs in != out");
  if ((m.scratch[0] && !s1) || (m.scratch[1] && !s2)) fail(ERR_NO_BUFFER, "scratch not set");
  if (prog.npasses() == 0) {
    hip_check(hipMemcpyAsync(out, in, (size_t)elems * (prec ? 16 : 8), hipMemcpyDeviceToDevice,
                             stream),
              "hipMemcpyAsync");
    return;
  }
  void* slots[SLOT_COUNT] = {const_cast<void*>(in), out, s1, s2};
  prog.run(m, slots, dir, stream);
}

void NDBackend::execute_preserve(const void* in, void* out, void* s1, void* s2, int dir,
                                 hipStream_t stream) const {
  run_preserve(prog_, map_pr_, elems_, prec_, in, out, s1, s2, dir, stream);
}

// Single-pass launch with the caller's buffer sizes checked against the
// pass's extents (intermediate passes of a Bluestein chain span the padded
// length, longer than the transform).
static void check_pass_buffers(const PassProgram& prog, size_t i, int64_t data_elems,
                               const void* in, int64_t in_elems, const void* out,
                               int64_t out_elems) {
  if (i >= prog.npasses()) fail(ERR_INVALID_DIM, "pass index out of range");
  if (!in || !out) fail(ERR_NO_BUFFER, "buffer not set");
  if (in_elems < prog.pass_in_elems(i, data_elems) || out_elems < prog.pass_out_elems(i))
    fail(ERR_INVALID_DIM, "pass " + std::to_string(i) + " needs buffers of " +
                              std::to_string(prog.pass_in_elems(i, data_elems)) + " / " +
                              std::to_string(prog.pass_out_elems(i)) + " elements");
}

void NDBackend::pass_extents(size_t i, int64_t* in_elems, int64_t* out_elems) const {
  if (i >= prog_.npasses()) fail(ERR_INVALID_DIM, "pass index out of range");
  *in_elems = prog_.pass_in_elems(i, elems_);
  *out_elems = prog_.pass_out_elems(i);
}

void NDBackend::execute_pass(size_t i, void* in, int64_t in_elems, void* out, int64_t out_elems,
                             int dir) {
  check_pass_buffers(prog_, i, elems_, in, in_elems, out, out_elems);
  PassProgram::SlotMap m;
  m.in.assign(prog_.npasses(), SLOT_DATA);
  m.out.assign(prog_.npasses(), SLOT_WORK);
  void* slots[SLOT_COUNT] = {in, out, s1_.p, s2_.p};
  prog_.run_one(m, i, slots, dir, stream_);
}

OneDBackend::OneDBackend(int64_t n, int prec, bool own_scratch)
    : n_(n), prec_(prec), prog_(prec) {
  if (prec != 0 && prec != 1) fail(ERR_INVALID_FFTTYPE, "invalid FFT type");
  if (n < 1) fail(ERR_INVALID_DIM, "invalid length");
  prog_.add_transform(n, std::vector<Digit>{}, 1, 1);
  map_pr_ = prog_.assign(PassProgram::PRESERVE, n_);
  map_ip_ = prog_.assign(PassProgram::INPLACE, n_);
  if (own_scratch) ensure_scratch(map_pr_, prec_, s1_, s2_);
  if (diag_enabled()) {
    std::ostringstream os;
    os << "N=" << n << " prec=" << (prec ? "Z2Z" : "C2C") << " passes=\n" << prog_.describe();
    diag("fft1dPlan", os.str());
  }
}

void OneDBackend::execute(void* in, void* out, int dir) {
  if (!in || !out) fail(ERR_NO_BUFFER, "buffer not set");
  const bool inplace = in == out;
  const PassProgram::SlotMap& m = inplace ? map_ip_ : map_pr_;
  ensure_scratch(m, prec_, s1_, s2_);
  if (prog_.npasses() == 0) {
    if (!inplace)
      hip_check(hipMemcpyAsync(out, in, (size_t)n_ * (prec_ ? 16 : 8), hipMemcpyDeviceToDevice,
                               stream_),
                "hipMemcpyAsync");
    return;
  }
  void* slots[SLOT_COUNT] = {in, out, s1_.p, s2_.p};
  prog_.run(m, slots, dir, stream_);
}

void OneDBackend::execute_in_scratch(const void* in, void* out, int dir, void* scratch,
                                     size_t scratch_bytes) const {
  if (!in || !out) fail(ERR_NO_BUFFER, "buffer not set");
  const bool inplace = in == out;
  const PassProgram::SlotMap& m = inplace ? map_ip_ : map_pr_;
  const size_t esz = prec_ ? 16 : 8;
  const size_t b0 = ((size_t)m.scratch[0] * esz + 255) & ~size_t(255);
  const size_t need = b0 + (size_t)m.scratch[1] * esz;
  if (need > scratch_bytes)
    fail(ERR_ALLOC, "FFT1D of length " + std::to_string(n_) + " needs " + std::to_string(need) +
                        " bytes of scratch; the allocation-free execute has " +
                        std::to_string(scratch_bytes));
  if (prog_.npasses() == 0) {
    if (!inplace)
      hip_check(hipMemcpyAsync(out, in, (size_t)n_ * esz, hipMemcpyDeviceToDevice, stream_),
                "hipMemcpyAsync");
    return;
  }
  char* base = static_cast<char*>(scratch);
  void* slots[SLOT_COUNT] = {const_cast<void*>(in), out, m.scratch[0] ? base : nullptr,
                             m.scratch[1] ? base + b0 : nullptr};
  prog_.run(m, slots, dir, stream_);
}

void OneDBackend::execute_preserve(const void* in, void* out, void* s1, void* s2, int dir,
                                   hipStream_t stream) const {
  run_preserve(prog_, map_pr_, n_, prec_, in, out, s1, s2, dir, stream);
}

void OneDBackend::pass_extents(size_t i, int64_t* in_elems, int64_t* out_elems) const {
  if (i >= prog_.npasses()) fail(ERR_INVALID_DIM, "pass index out of range");
  *in_elems = prog_.pass_in_elems(i, n_);
  *out_elems = prog_.pass_out_elems(i);
}

void OneDBackend::execute_pass(size_t i, void* in, int64_t in_elems, void* out, int64_t out_elems,
                               int dir) {
  check_pass_buffers(prog_, i, n_, in, in_elems, out, out_elems);
  PassProgram::SlotMap m;
  m.in.assign(prog_.npasses(), SLOT_DATA);
  m.out.assign(prog_.npasses(), SLOT_WORK);
  void* slots[SLOT_COUNT] = {in, out, s1_.p, s2_.p};
  prog_.run_one(m, i, slots, dir, stream_);
}

void scale_async(void* data, double f, uint64_t count, int prec, hipStream_t stream) {
  hip_check((hipError_t)sa_launch_scale(data, f, count, prec, stream), "normalize launch");
}

}  // namespace sa
