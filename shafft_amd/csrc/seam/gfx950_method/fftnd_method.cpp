// gfx950 backend of the local N-D transform: each seam function forwards to
// the library's C ABI (include/shafft_amd.h section 1).  Statuses are the
// library's shafft_status_t values (the same integers as shafft::Status);
// the core wraps non-zero returns in ERR_BACKEND (SHAFFT_BACKEND_CHECK,
// src/cpp/common/error_macros.hpp:57-79) and the library's thread-local last
// error keeps the detail.
#include "fftnd_method.hpp"

int fftndPlan(FFTNDHandle& plan, int nta, int* ta, int ndim, int* dimensions,
              shafft::FFTType precision, void* in, void* out) {
  if (plan.plan) {
    shafftAmdNDDestroy(plan.plan);
    plan.plan = nullptr;
  }
  const int prec = precision == shafft::FFTType::Z2Z ? 1 : 0;
  const int rc = shafftAmdNDPlan(&plan.plan, nta, ta, ndim, dimensions, prec, in, out);
  if (rc == 0 && plan.stream) return shafftAmdNDSetStream(plan.plan, plan.stream);
  return rc;
}

int fftndSetStream(FFTNDHandle& plan, hipStream_t stream) {
  plan.stream = stream;
  return shafftAmdNDSetStream(plan.plan, stream);
}

// By value, as the core calls it (shafft_internal.cpp:94): the plan pointer
// is shared with the caller's copy; data/work come back swapped when the
// result ends in the work buffer, after the stream has drained.
int fftndExecute(FFTNDHandle plan, void*& data, void*& work, shafft::FFTDirection direction) {
  return shafftAmdNDExecute(plan.plan, &data, &work,
                            direction == shafft::FFTDirection::FORWARD ? 0 : 1);
}

int fftndDestroy(FFTNDHandle plan) {
  return plan.plan ? shafftAmdNDDestroy(plan.plan) : 0;
}
