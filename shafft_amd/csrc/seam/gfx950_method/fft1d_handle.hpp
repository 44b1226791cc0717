// Handle and configuration of the distributed 1-D FFT, gfx950 backend.  The
// fields of src/cpp/fft/hipfft_method/fft1d_handle.hpp:14-44 (the core holds
// a pointer to the handle, plan_1d_data.hpp:23-25, and passes it back to the
// seam); `internal` holds the library's shafft_amd_1d_plan.
#ifndef SHAFFT_FFT1D_HANDLE_HPP
#define SHAFFT_FFT1D_HANDLE_HPP

#include <shafft/shafft_types.hpp>

#include <cstddef>
#include <hip/hip_runtime_api.h>
#include <mpi.h>

namespace shafft {

/// Path selection options (kept for the interface: at one rank the library
/// picks the four-step or Bluestein path from the length alone, and never
/// pads, so the transform is always the exact N-point DFT).
struct FFT1DConfig {
  double tau = 1.25;
  bool allowPadding = false;
};

struct FFT1DHandle {
  void* internal = nullptr;  ///< shafft_amd_1d_plan*

  size_t globalN = 0;

  size_t localNInit = 0;
  size_t localStartInit = 0;

  size_t localNTrans = 0;
  size_t localStartTrans = 0;

  size_t localAllocSize = 0;

  FFTType precision = FFTType::C2C;
  MPI_Comm comm = MPI_COMM_NULL;
  hipStream_t stream = nullptr;
};

}  // namespace shafft

#endif  // SHAFFT_FFT1D_HANDLE_HPP
