// The part of SHAFFT's include/shafft/shafft_types.hpp and shafft_enums.hpp
// that the backend seam headers use: the FFTType and FFTDirection
// enumerations with the reference's enumerator order (an interface the
// adapter must match).  Used only to compile the gfx950 seam adapter and its
// conformance program outside SHAFFT's tree; inside it the core's own header
// is found first.
#ifndef SHAFFT_TYPES_H
#define SHAFFT_TYPES_H

#include <mpi.h>

namespace shafft {
enum class FFTDirection { FORWARD, BACKWARD };
enum class FFTType { C2C, Z2Z };
}  // namespace shafft

#endif  // SHAFFT_TYPES_H
