// Normalisation seam of the gfx950 backend over the library's scale kernel.
#include "normalize.hpp"

#include <shafft_amd.h>

int normalizeComplexFloat(float norm_factor, size_t tensor_size, void* data, hipStream_t stream) {
  return shafftAmdNormalizeComplexFloat(norm_factor, tensor_size, data, stream);
}

int normalizeComplexDouble(double norm_factor, size_t tensor_size, void* data,
                           hipStream_t stream) {
  return shafftAmdNormalizeComplexDouble(norm_factor, tensor_size, data, stream);
}
