// Twiddle, four-step and Bluestein tables.  Every root of unity is
// exp(-2 pi i m / L) for an integer m reduced modulo L in exact integer
// arithmetic, evaluated in long double and rounded once to the working
// precision.  The Bluestein chirp index is n^2 mod 2N in 64-bit integers
// (the reference evaluates pi n^2 / N in double without the reduction:
// src/cpp/fft/hipfft_method/1d/fft1d_bluestein_executor.hip:150-205).
#pragma once

#include <complex>
#include <cstdint>
#include <vector>

namespace sa {

using cld = std::complex<long double>;

// exp(-2 pi i m / L)
cld root_fwd(int64_t m, int64_t L);

// Stage twiddles of a Stockham plan: for stage s (Ns = product of the
// earlier radices, R = radix) the entries W_{Ns R}^{rem r}, rem < Ns,
// 1 <= r < R, laid out as [rem][r-1].  `offsets` receives the start of each
// stage's block.
std::vector<cld> stage_twiddles(const std::vector<int>& radices, std::vector<int>& offsets);

// Two-level table for W_N^e, e < N: lo[j] = W^j (j < 2^shift),
// hi[i] = W^{i 2^shift}.
void two_level(int64_t N, int& shift, std::vector<cld>& lo, std::vector<cld>& hi);

// Bluestein tables for a length-n DFT through a length-M (power of two,
// M >= 2n-1) cyclic convolution, for direction sign (-1 forward, +1 back):
//   chirp[k] = exp(sign i pi (k^2 mod 2n) / n),            k < n
//   H        = FFT_M(h) / M,  h[m] = conj(chirp[|m|]) for |m| < n (cyclic)
void bluestein_tables(int64_t n, int64_t M, int sign, std::vector<cld>& chirp,
                      std::vector<cld>& H);

// Both directions at once ([0] forward, [1] backward).  h is even
// (h[m] = h[M-m]), so its transform is even and the backward tables are the
// conjugates of the forward ones: chirp_b = conj(chirp_f) and
// H_b[k] = conj(H_f[-k]) = conj(H_f[k]).  One transform instead of two.
void bluestein_tables_pair(int64_t n, int64_t M, std::vector<cld> chirp[2], std::vector<cld> H[2]);

// In-place radix-2 FFT in long double (forward sign), M a power of two;
// large transforms run on up to 16 host threads.
void fft_ld(std::vector<cld>& a);

}  // namespace sa
