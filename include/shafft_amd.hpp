/*
 * C++ class API over libshafft_amd.so, mirroring the reference's
 * include/shafft/shafft.hpp (namespace shafft: FFTND, FFT1D, buffer helpers,
 * configuration queries) at one rank.  Header-only: every call forwards to
 * the C ABI in shafft_amd.h and returns its status code (0 on success), so a
 * program written against the reference's classes compiles against this
 * header with the MPI_Comm arguments replaced by shafft::kOneRank (0).
 *
 * Semantics that differ from the reference's classes are the C ABI's: one
 * rank, process grid {1,...,1}, nothing distributed.
 */
#ifndef SHAFFT_AMD_HPP
#define SHAFFT_AMD_HPP

#include <climits>
#include <complex>
#include <cstddef>
#include <utility>
#include <vector>

#include "shafft_amd.h"

namespace shafft {

using complexf = std::complex<float>;   // shafft_types.hpp:71
using complexd = std::complex<double>;  // shafft_types.hpp:80
using Comm = shafft_comm_t;             // stands in for MPI_Comm
constexpr Comm kOneRank = 0;

enum class FFTDirection { FORWARD, BACKWARD };                  // shafft_enums.hpp:18
enum class FFTType { C2C, Z2Z };                                // shafft_enums.hpp:30
enum class TensorLayout { CURRENT, INITIAL, REDISTRIBUTED };    // shafft_enums.hpp:42
enum class TransformLayout { REDISTRIBUTED, INITIAL };          // shafft_types.hpp:128
enum class DecompositionStrategy { MAXIMIZE_NDA, MINIMIZE_NDA };  // shafft_types.hpp:114

struct Version {
  int major = 0, minor = 0, patch = 0;
};

inline const char* getBackendName() noexcept { return shafftGetBackendName(); }
inline Version getVersion() noexcept {
  Version v;
  shafftGetVersion(&v.major, &v.minor, &v.patch);
  return v;
}
inline const char* getVersionString() noexcept { return shafftGetVersionString(); }
[[nodiscard]] inline int finalize() noexcept { return shafftFinalize(); }

// Device buffers (shafft.hpp:760-801).
[[nodiscard]] inline int allocBuffer(size_t count, complexf** buf) noexcept {
  return shafftAllocBufferF(count, reinterpret_cast<void**>(buf));
}
[[nodiscard]] inline int allocBuffer(size_t count, complexd** buf) noexcept {
  return shafftAllocBufferD(count, reinterpret_cast<void**>(buf));
}
[[nodiscard]] inline int freeBuffer(complexf* buf) noexcept { return shafftFreeBufferF(buf); }
[[nodiscard]] inline int freeBuffer(complexd* buf) noexcept { return shafftFreeBufferD(buf); }
[[nodiscard]] inline int copyToBuffer(complexf* dst, const complexf* src, size_t count) noexcept {
  return shafftCopyToBufferF(dst, src, count);
}
[[nodiscard]] inline int copyToBuffer(complexd* dst, const complexd* src, size_t count) noexcept {
  return shafftCopyToBufferD(dst, src, count);
}
[[nodiscard]] inline int copyFromBuffer(complexf* dst, const complexf* src, size_t count) noexcept {
  return shafftCopyFromBufferF(dst, src, count);
}
[[nodiscard]] inline int copyFromBuffer(complexd* dst, const complexd* src, size_t count) noexcept {
  return shafftCopyFromBufferD(dst, src, count);
}

namespace detail {
inline shafft_t c_type(FFTType t) { return t == FFTType::Z2Z ? SHAFFT_Z2Z : SHAFFT_C2C; }
inline shafft_tensor_layout_t c_layout(TensorLayout l) {
  return l == TensorLayout::INITIAL         ? SHAFFT_TENSOR_LAYOUT_INITIAL
         : l == TensorLayout::REDISTRIBUTED ? SHAFFT_TENSOR_LAYOUT_REDISTRIBUTED
                                            : SHAFFT_TENSOR_LAYOUT_CURRENT;
}

// Plan handle shared by FFTND and FFT1D: created on init, destroyed on
// release; moves transfer it.
class PlanHandle {
 public:
  PlanHandle() noexcept = default;
  ~PlanHandle() { release(); }
  PlanHandle(PlanHandle&& o) noexcept : h_(std::exchange(o.h_, nullptr)) {}
  PlanHandle& operator=(PlanHandle&& o) noexcept {
    if (this != &o) {
      release();
      h_ = std::exchange(o.h_, nullptr);
    }
    return *this;
  }
  PlanHandle(const PlanHandle&) = delete;
  PlanHandle& operator=(const PlanHandle&) = delete;
  void release() noexcept {
    if (h_) (void)shafftDestroy(&h_);
    h_ = nullptr;
  }
  void* get() const noexcept { return h_; }
  void** addr() noexcept { return &h_; }

 private:
  void* h_ = nullptr;
};
}  // namespace detail

// Common plan operations (the reference's FFT base class).
class FFT {
 public:
  virtual ~FFT() = default;
  [[nodiscard]] int plan() noexcept { return shafftPlan(h_.get()); }
  [[nodiscard]] int execute(FFTDirection d) noexcept {
    return shafftExecute(h_.get(), d == FFTDirection::FORWARD ? SHAFFT_FORWARD : SHAFFT_BACKWARD);
  }
  [[nodiscard]] int normalize() noexcept { return shafftNormalize(h_.get()); }
  [[nodiscard]] int setBuffers(complexf* data, complexf* work) noexcept { return setBuffersRaw(data, work); }
  [[nodiscard]] int setBuffers(complexd* data, complexd* work) noexcept { return setBuffersRaw(data, work); }
  [[nodiscard]] int getBuffers(complexf** data, complexf** work) noexcept {
    return getBuffersRaw(reinterpret_cast<void**>(data), reinterpret_cast<void**>(work));
  }
  [[nodiscard]] int getBuffers(complexd** data, complexd** work) noexcept {
    return getBuffersRaw(reinterpret_cast<void**>(data), reinterpret_cast<void**>(work));
  }
  [[nodiscard]] int setBuffersRaw(void* data, void* work) noexcept {
    return shafftSetBuffers(h_.get(), data, work);
  }
  [[nodiscard]] int getBuffersRaw(void** data, void** work) noexcept {
    return shafftGetBuffers(h_.get(), data, work);
  }
  [[nodiscard]] int setStream(void* hipStream) noexcept { return shafftSetStream(h_.get(), hipStream); }
  [[nodiscard]] size_t allocSize() const noexcept {
    size_t n = 0;
    return shafftGetAllocSize(h_.get(), &n) == 0 ? n : 0;
  }
  [[nodiscard]] size_t globalSize() const noexcept {
    size_t n = 0;
    return shafftGetGlobalSize(h_.get(), &n) == 0 ? n : 0;
  }
  [[nodiscard]] bool isActive() const noexcept {
    int a = 0;
    return shafftIsActive(h_.get(), &a) == 0 && a != 0;
  }
  [[nodiscard]] FFTType fftType() const noexcept { return type_; }
  [[nodiscard]] int getCommunicator(Comm* out) const noexcept { return shafftGetCommunicator(h_.get(), out); }
  virtual void release() noexcept { h_.release(); }
  void* handle() const noexcept { return h_.get(); }

 protected:
  FFT() noexcept = default;
  FFT(FFT&&) noexcept = default;
  FFT& operator=(FFT&&) noexcept = default;
  detail::PlanHandle h_;
  FFTType type_ = FFTType::C2C;
};

// N-D transform (shafft.hpp:51-324).
class FFTND : public FFT {
 public:
  FFTND() noexcept = default;
  FFTND(FFTND&&) noexcept = default;
  FFTND& operator=(FFTND&&) noexcept = default;

  [[nodiscard]] int init(const std::vector<int>& commDims, const std::vector<size_t>& dimensions,
                         FFTType type, Comm comm,
                         TransformLayout output = TransformLayout::REDISTRIBUTED) noexcept {
    if (commDims.size() != dimensions.size()) return SHAFFT_ERR_DIM_MISMATCH;
    std::vector<int> dims(dimensions.size()), grid(commDims);
    for (size_t i = 0; i < dimensions.size(); ++i) {
      if (dimensions[i] > (size_t)INT_MAX) return SHAFFT_ERR_SIZE_OVERFLOW;
      dims[i] = (int)dimensions[i];
    }
    const int rc = create();
    if (rc) return rc;
    type_ = type;
    ndim_ = (int)dims.size();
    return shafftNDInit(h_.get(), (int)dims.size(), grid.data(), dims.data(), detail::c_type(type), comm,
                        output == TransformLayout::INITIAL ? SHAFFT_LAYOUT_INITIAL
                                                           : SHAFFT_LAYOUT_REDISTRIBUTED);
  }
  [[nodiscard]] int init(shafft_nd_config_t& cfg) noexcept {
    const int rc = create();
    if (rc) return rc;
    type_ = cfg.precision == SHAFFT_Z2Z ? FFTType::Z2Z : FFTType::C2C;
    ndim_ = cfg.ndim;
    return shafftNDInitFromConfig(h_.get(), &cfg);
  }
  [[nodiscard]] int ndim() const noexcept { return ndim_; }
  [[nodiscard]] int getLayout(std::vector<size_t>& subsize, std::vector<size_t>& offset,
                              TensorLayout layout = TensorLayout::CURRENT) const noexcept {
    subsize.assign((size_t)ndim_, 0);
    offset.assign((size_t)ndim_, 0);
    return shafftGetLayout(h_.get(), subsize.data(), offset.data(), detail::c_layout(layout));
  }
  [[nodiscard]] int getAxes(std::vector<int>& ca, std::vector<int>& da,
                            TensorLayout layout = TensorLayout::CURRENT) const noexcept {
    ca.assign((size_t)ndim_, 0);
    da.clear();  // one rank: nothing distributed
    return shafftGetAxes(h_.get(), ca.data(), nullptr, detail::c_layout(layout));
  }

 private:
  int create() noexcept {
    h_.release();
    return shafftNDCreate(h_.addr());
  }
  int ndim_ = 0;
};

// 1-D transform (shafft.hpp:341-627).
class FFT1D : public FFT {
 public:
  FFT1D() noexcept = default;
  FFT1D(FFT1D&&) noexcept = default;
  FFT1D& operator=(FFT1D&&) noexcept = default;

  [[nodiscard]] int init(size_t globalN, size_t localN, size_t localStart, FFTType type,
                         Comm comm) noexcept {
    h_.release();
    const int rc = shafft1DCreate(h_.addr());
    if (rc) return rc;
    type_ = type;
    return shafft1DInit(h_.get(), globalN, localN, localStart, detail::c_type(type), comm);
  }
  [[nodiscard]] int init(shafft_1d_config_t& cfg) noexcept {
    h_.release();
    const int rc = shafft1DCreate(h_.addr());
    if (rc) return rc;
    type_ = cfg.precision == SHAFFT_Z2Z ? FFTType::Z2Z : FFTType::C2C;
    return shafft1DInitFromConfig(h_.get(), &cfg);
  }
  [[nodiscard]] int ndim() const noexcept { return 1; }
  [[nodiscard]] int getLayout(std::vector<size_t>& localShape, std::vector<size_t>& offset,
                              TensorLayout layout = TensorLayout::CURRENT) const noexcept {
    localShape.assign(1, 0);
    offset.assign(1, 0);
    return shafftGetLayout(h_.get(), localShape.data(), offset.data(), detail::c_layout(layout));
  }
  // ca = {}, da = {0}: the single axis is the (block-)distributed one.
  [[nodiscard]] int getAxes(std::vector<int>& ca, std::vector<int>& da,
                            TensorLayout layout = TensorLayout::CURRENT) const noexcept {
    ca.clear();
    da.assign(1, 0);
    return shafftGetAxes(h_.get(), nullptr, da.data(), detail::c_layout(layout));
  }
};

// Configuration queries (shafft.hpp:843-880).
[[nodiscard]] inline int configurationND(const std::vector<size_t>& size, FFTType precision,
                                         std::vector<int>& commDims, int& nda,
                                         std::vector<size_t>& subsize, std::vector<size_t>& offset,
                                         int& commSize, DecompositionStrategy strategy,
                                         size_t memLimit, Comm comm) noexcept {
  const int nd = (int)size.size();
  std::vector<int> dims((size_t)nd);
  for (int i = 0; i < nd; ++i) {
    if (size[(size_t)i] > (size_t)INT_MAX) return SHAFFT_ERR_SIZE_OVERFLOW;
    dims[(size_t)i] = (int)size[(size_t)i];
  }
  commDims.resize((size_t)nd, 0);
  subsize.assign((size_t)nd, 0);
  offset.assign((size_t)nd, 0);
  return shafftConfigurationND(nd, dims.data(), detail::c_type(precision), commDims.data(), &nda,
                               subsize.data(), offset.data(), &commSize,
                               strategy == DecompositionStrategy::MINIMIZE_NDA ? SHAFFT_MINIMIZE_NDA
                                                                               : SHAFFT_MAXIMIZE_NDA,
                               memLimit, comm);
}
[[nodiscard]] inline int configuration1D(size_t globalN, size_t& localN, size_t& localStart,
                                         FFTType precision, Comm comm) noexcept {
  size_t alloc = 0;
  return shafftConfiguration1D(globalN, &localN, &localStart, &alloc, detail::c_type(precision), comm);
}

}  // namespace shafft

#endif /* SHAFFT_AMD_HPP */
