// Backend seam of the distributed 1-D FFT, gfx950 backend: the declarations
// of src/cpp/fft/hipfft_method/fft1d_method.hpp:14-43, without the hipFFT
// include.
#ifndef SHAFFT_FFT1D_METHOD_HPP
#define SHAFFT_FFT1D_METHOD_HPP

#include "fft1d_handle.hpp"

namespace shafft {

int fft1dPlan(FFT1DHandle& handle,
              size_t globalN,
              FFTType precision,
              MPI_Comm comm,
              void* in,
              void* out,
              const FFT1DConfig& config = FFT1DConfig{});

int fft1dExecute(FFT1DHandle& handle, void* in, void* out, FFTDirection direction);

int fft1dDestroy(FFT1DHandle& handle);

int fft1dSetStream(FFT1DHandle& handle, hipStream_t stream);

int fft1dSynchronize(FFT1DHandle& handle);

int fft1dNormalize(FFT1DHandle& handle, void* data, int normExponent, size_t localCount);

int fft1dQueryLayout(size_t globalN,
                     size_t& localN,
                     size_t& localStart,
                     size_t& localAllocSize,
                     size_t& localNTrans,
                     size_t& localStartTrans,
                     MPI_Comm comm,
                     const FFT1DConfig& config = FFT1DConfig{});

inline int fft1dFinalize() noexcept {
  return 0;
}

}  // namespace shafft

#endif  // SHAFFT_FFT1D_METHOD_HPP
