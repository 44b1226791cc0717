// Backend seam of the local N-D transform, gfx950 backend.  The same
// declarations as src/cpp/fft/hipfft_method/fftnd_method.hpp:14-31, without
// the hipFFT include: the core calls these for every subplan of a slab stage
// (the slab exchange between stages stays in the core).
#ifndef FFTND_METHOD_HPP
#define FFTND_METHOD_HPP

#include <shafft/shafft_types.hpp>

#include "fftnd_handle.hpp"

int fftndPlan(FFTNDHandle& plan,
              int nta,
              int* ta,
              int ndim,
              int* dimensions,
              shafft::FFTType precision,
              void* in,
              void* out);

int fftndSetStream(FFTNDHandle& plan, hipStream_t stream);

int fftndExecute(FFTNDHandle plan, void*& data, void*& work, shafft::FFTDirection direction);

int fftndDestroy(FFTNDHandle plan);

inline int fftndFinalize() noexcept {
  return shafftAmdNDFinalize();
}

#endif  // FFTND_METHOD_HPP
