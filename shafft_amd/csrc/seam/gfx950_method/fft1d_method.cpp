// gfx950 backend of the 1-D FFT seam over the library's C ABI.  One rank per
// communicator: the library runs the whole N-point transform on one GPU
// (four-step for 7-smooth N, Bluestein otherwise, exact N-point DFT), so a
// communicator of more than one rank is refused with ERR_NOT_IMPL and the
// layout is the one-rank layout (localN = N, localStart = 0).
#include "fft1d_method.hpp"

#include <shafft_amd.h>

namespace shafft {

namespace {

shafft_amd_1d_plan* plan_of(const FFT1DHandle& h) {
  return static_cast<shafft_amd_1d_plan*>(h.internal);
}

// SHAFFT_ERR_INVALID_COMM (2) for a null communicator, SHAFFT_ERR_NOT_IMPL
// (14) for more than one rank, 0 at one rank.
int one_rank(MPI_Comm comm) {
  if (comm == MPI_COMM_NULL) return SHAFFT_ERR_INVALID_COMM;
  int size = 0;
  if (MPI_Comm_size(comm, &size) != MPI_SUCCESS) return SHAFFT_ERR_INVALID_COMM;
  return size == 1 ? 0 : SHAFFT_ERR_NOT_IMPL;
}

}  // namespace

int fft1dQueryLayout(size_t globalN, size_t& localN, size_t& localStart, size_t& localAllocSize,
                     size_t& localNTrans, size_t& localStartTrans, MPI_Comm comm,
                     const FFT1DConfig&) {
  if (const int rc = one_rank(comm)) return rc;
  return shafftAmd1DQueryLayout(globalN, &localN, &localStart, &localAllocSize, &localNTrans,
                                &localStartTrans);
}

int fft1dPlan(FFT1DHandle& handle, size_t globalN, FFTType precision, MPI_Comm comm, void* in,
              void* out, const FFT1DConfig& config) {
  if (const int rc = one_rank(comm)) return rc;
  if (handle.internal) fft1dDestroy(handle);
  shafft_amd_1d_plan* p = nullptr;
  int rc = shafftAmd1DPlan(&p, globalN, precision == FFTType::Z2Z ? 1 : 0, in, out);
  if (rc) return rc;
  handle.internal = p;
  handle.globalN = globalN;
  handle.precision = precision;
  handle.comm = comm;
  rc = fft1dQueryLayout(globalN, handle.localNInit, handle.localStartInit, handle.localAllocSize,
                        handle.localNTrans, handle.localStartTrans, comm, config);
  if (rc == 0 && handle.stream) rc = shafftAmd1DSetStream(p, handle.stream);
  return rc;
}

// in -> out (in == out allowed), then waits for the stream: the core swaps
// its buffers right after the call and treats the result as ready
// (src/cpp/shafft_internal.cpp:584-611).
int fft1dExecute(FFT1DHandle& handle, void* in, void* out, FFTDirection direction) {
  if (!handle.internal) return SHAFFT_ERR_PLAN_NOT_INIT;
  const int rc = shafftAmd1DExecute(plan_of(handle), in, out,
                                    direction == FFTDirection::FORWARD ? 0 : 1);
  return rc ? rc : shafftAmd1DSynchronize(plan_of(handle));
}

int fft1dDestroy(FFT1DHandle& handle) {
  const int rc = handle.internal ? shafftAmd1DDestroy(plan_of(handle)) : 0;
  handle.internal = nullptr;
  return rc;
}

int fft1dSetStream(FFT1DHandle& handle, hipStream_t stream) {
  if (!handle.internal) return SHAFFT_ERR_PLAN_NOT_INIT;
  handle.stream = stream;
  return shafftAmd1DSetStream(plan_of(handle), stream);
}

int fft1dSynchronize(FFT1DHandle& handle) {
  if (!handle.internal) return SHAFFT_ERR_PLAN_NOT_INIT;
  return shafftAmd1DSynchronize(plan_of(handle));
}

// Factor 1/sqrt(N)^normExponent in the working precision, then waits
// (fft1d_method.hip:652-679 of the hipFFT backend).
int fft1dNormalize(FFT1DHandle& handle, void* data, int normExponent, size_t localCount) {
  if (!handle.internal) return SHAFFT_ERR_PLAN_NOT_INIT;
  return shafftAmd1DNormalize(plan_of(handle), data, normExponent, localCount);
}

}  // namespace shafft
