// Per-subplan backend state of the gfx950 backend.  Replaces
// src/cpp/fft/hipfft_method/fftnd_handle.hpp:26-47: the SHAFFT core
// allocates it with `new FFTNDHandle[n]()` and sets and reads its `stream`
// and `comm` members (src/cpp/shafft_internal.cpp:149,351,368); everything
// else the hipFFT handle held (subplans, superbatch loops, gpuTT transpose
// plans) lives inside the library's opaque plan.
#ifndef FFTND_HANDLE_HPP
#define FFTND_HANDLE_HPP

#include <hip/hip_runtime_api.h>
#include <mpi.h>
#include <shafft_amd.h>

struct FFTNDHandle {
  shafft_amd_nd_plan* plan = nullptr;  // owned, released by fftndDestroy
  hipStream_t stream = nullptr;
  MPI_Comm comm = MPI_COMM_NULL;
};

#endif  // FFTND_HANDLE_HPP
