// Normalisation seam, gfx950 backend: the declarations of
// src/cpp/fft/hipfft_method/common/normalize.hpp.
#ifndef SHAFFT_NORMALIZE_HPP
#define SHAFFT_NORMALIZE_HPP

#include <cstddef>
#include <hip/hip_runtime_api.h>

/// x[i] *= norm_factor over tensor_size complex floats (asynchronous).
int normalizeComplexFloat(float norm_factor,
                          size_t tensor_size,
                          void* data,
                          hipStream_t stream = nullptr);

/// x[i] *= norm_factor over tensor_size complex doubles (asynchronous).
int normalizeComplexDouble(double norm_factor,
                           size_t tensor_size,
                           void* data,
                           hipStream_t stream = nullptr);

#endif  // SHAFFT_NORMALIZE_HPP
