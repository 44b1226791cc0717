"""The kernels' in-register butterflies (shafft_amd/csrc/kernels/fft_device.hpp:
radices 2..28 incl. the prime-factor and Cooley-Tukey composite forms) built
for the host with g++ and checked against the direct DFT, both directions,
float and double.  The header's one device intrinsic, __umulhi, gets its host
definition (the high 32 bits of the 64-bit product)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RADICES = (2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 14, 15, 16, 18, 20, 21, 24, 25, 28)

SRC = r"""
#include <cstdint>
#include <cstdio>
static inline uint32_t __umulhi(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a * b) >> 32); }
#include "shafft_amd/csrc/kernels/fft_device.hpp"

template <int RAD, int SIGN, typename R>
void run(const double* in, double* out) {
  sa::cpx<R> v[RAD];
  for (int i = 0; i < RAD; ++i) v[i] = {(R)in[2 * i], (R)in[2 * i + 1]};
  sa::dft<RAD, SIGN>(v);
  for (int i = 0; i < RAD; ++i) { out[2 * i] = v[i].x; out[2 * i + 1] = v[i].y; }
}
template <int RAD>
void all(const double* in, double* out) {  // out: [fwd f, bwd f, fwd d, bwd d]
  run<RAD, -1, float>(in, out);
  run<RAD, +1, float>(in, out + 2 * RAD);
  run<RAD, -1, double>(in, out + 4 * RAD);
  run<RAD, +1, double>(in, out + 6 * RAD);
}
int main() {
  double in[64], out[256];
  int rad;
  while (std::scanf("%d", &rad) == 1) {
    for (int i = 0; i < 2 * rad; ++i) std::scanf("%lf", &in[i]);
    switch (rad) {
      case 2: all<2>(in, out); break;   case 3: all<3>(in, out); break;
      case 4: all<4>(in, out); break;   case 5: all<5>(in, out); break;
      case 6: all<6>(in, out); break;   case 7: all<7>(in, out); break;
      case 8: all<8>(in, out); break;   case 9: all<9>(in, out); break;
      case 10: all<10>(in, out); break; case 12: all<12>(in, out); break;
      case 16: all<16>(in, out); break; case 14: all<14>(in, out); break;
      case 15: all<15>(in, out); break; case 18: all<18>(in, out); break;
      case 20: all<20>(in, out); break; case 21: all<21>(in, out); break;
      case 24: all<24>(in, out); break; case 25: all<25>(in, out); break;
      case 28: all<28>(in, out); break; default: return 2;
    }
    for (int i = 0; i < 8 * rad; ++i) std::printf("%.17g\n", out[i]);
  }
  return 0;
}
"""


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    d = tmp_path_factory.mktemp("butterflies")
    src, exe = d / "b.cpp", d / "b"
    src.write_text(SRC)
    subprocess.run(["g++", "-O1", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    f"-I{ROOT}", str(src), "-o", str(exe)], check=True, capture_output=True, text=True)
    return exe


def test_butterflies_match_direct_dft(harness):
    rng = np.random.default_rng(11)
    cases = [(r, rng.uniform(-1, 1, r) + 1j * rng.uniform(-1, 1, r)) for r in RADICES for _ in range(3)]
    cases += [(r, np.eye(r)[k].astype(complex)) for r in RADICES for k in range(r)]  # unit impulses
    stdin = "".join(f"{r} " + " ".join(f"{z.real:.17g} {z.imag:.17g}" for z in x) + "\n" for r, x in cases)
    res = subprocess.run([str(harness)], input=stdin, capture_output=True, text=True, check=True)
    vals = np.array(res.stdout.split(), dtype=float)
    pos = 0
    for r, x in cases:
        out = vals[pos:pos + 8 * r].reshape(4, r, 2)
        pos += 8 * r
        y = out[..., 0] + 1j * out[..., 1]
        n = np.arange(r)
        m = np.outer(n, n) % r  # exact index reduction before the angle
        fwd = np.exp(-2j * np.pi * m / r) @ x
        bwd = np.exp(2j * np.pi * m / r) @ x
        for got, ref, tol in ((y[0], fwd, 1e-6), (y[1], bwd, 1e-6), (y[2], fwd, 1e-14), (y[3], bwd, 1e-14)):
            assert np.max(np.abs(got - ref)) < tol * max(1.0, np.max(np.abs(ref))), (r, x)
    assert pos == vals.size
