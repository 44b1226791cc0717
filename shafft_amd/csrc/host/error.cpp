// Thread-local last-error record (mirrors the reference's error slots,
// src/cpp/shafft_error.cpp:9-11,159) and the SHAFFT_FFT_DIAG trace.
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>

#include "common.hpp"

namespace sa {

namespace {
thread_local int t_status = 0;
thread_local int t_source = 0;
thread_local int t_domain = 0;
thread_local std::string t_message;
}  // namespace

void set_last_error(int status, int source, int domain, const char* msg) {
  t_status = status;
  t_source = source;
  t_domain = domain;
  t_message = msg ? msg : "";
  if (diag_enabled()) diag("error", t_message);
}

void clear_last_error() {
  t_status = 0;
  t_source = 0;
  t_domain = 0;
  t_message.clear();
}

int last_error_status() { return t_status; }
int last_error_source() { return t_source; }
int last_error_domain() { return t_domain; }
const std::string& last_error_message() { return t_message; }

bool diag_enabled() {
  static int cached = -1;
  if (cached < 0) {
    const char* e = std::getenv("SHAFFT_FFT_DIAG");
    cached = (e && e[0] == '1') ? 1 : 0;
  }
  return cached != 0;
}

void diag(const char* where, const std::string& msg) {
  if (!diag_enabled()) return;
  std::cerr << "[SHAFFT:gfx950][r0][" << where << "] " << msg << std::endl;
}

}  // namespace sa
