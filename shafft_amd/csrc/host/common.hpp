// Host-side common definitions: status codes, error transport, HIP checks.
//
// Status values and error sources follow SHAFFT's own numbering
// (include/shafft/shafft_types.hpp:92-110, include/shafft/shafft.h:104-140)
// so a caller of the reference C API sees the same integers.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace sa {

enum Status : int {
  SUCCESS = 0,
  ERR_NULLPTR = 1,
  ERR_INVALID_COMM = 2,
  ERR_NO_BUFFER = 3,
  ERR_PLAN_NOT_INIT = 4,
  ERR_INVALID_DIM = 5,
  ERR_DIM_MISMATCH = 6,
  ERR_INVALID_DECOMP = 7,
  ERR_INVALID_FFTTYPE = 8,
  ERR_ALLOC = 9,
  ERR_BACKEND = 10,
  ERR_MPI = 11,
  ERR_INVALID_LAYOUT = 12,
  ERR_SIZE_OVERFLOW = 13,
  ERR_NOT_IMPL = 14,
  ERR_INVALID_STATE = 15,
  ERR_INTERNAL = 16,
};

// Error source domains (include/shafft/shafft.h: shafft_errsrc_t).
enum ErrSource : int {
  SRC_NONE = 0,
  SRC_MPI = 1,
  SRC_HIP = 2,
  SRC_HIPFFT = 3,
  SRC_FFTW = 4,
  SRC_SYSTEM = 5,
};

struct Error : std::runtime_error {
  int status;
  int source;
  int domain;
  Error(int st, int src, int dom, const std::string& msg)
      : std::runtime_error(msg), status(st), source(src), domain(dom) {}
};

// Thread-local last-error slot (one plan per thread, as in SHAFFT).
void set_last_error(int status, int source, int domain, const char* msg);
void clear_last_error();
int last_error_status();
int last_error_source();
int last_error_domain();
const std::string& last_error_message();

// Diagnostic trace, enabled by SHAFFT_FFT_DIAG=1 (same switch and tag style
// as src/cpp/fft/common/diag.hpp:19-60).
bool diag_enabled();
void diag(const char* where, const std::string& msg);

[[noreturn]] inline void fail(int status, const std::string& msg) {
  throw Error(status, SRC_NONE, 0, msg);
}

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    throw Error(ERR_BACKEND, SRC_HIP, static_cast<int>(e),
                std::string(what) + ": " + hipGetErrorString(e));
  }
}

// Runs f and converts exceptions into a status + last-error record.
template <typename F>
int guarded(F&& f) noexcept {
  try {
    clear_last_error();
    int rc = f();
    return rc;
  } catch (const Error& e) {
    set_last_error(e.status, e.source, e.domain, e.what());
    return e.status;
  } catch (const std::bad_alloc& e) {
    set_last_error(ERR_ALLOC, SRC_SYSTEM, 0, e.what());
    return ERR_ALLOC;
  } catch (const std::exception& e) {
    set_last_error(ERR_INTERNAL, SRC_NONE, 0, e.what());
    return ERR_INTERNAL;
  } catch (...) {
    set_last_error(ERR_INTERNAL, SRC_NONE, 0, "unknown exception");
    return ERR_INTERNAL;
  }
}

}  // namespace sa
