"""The multiply-high divisor used by the kernels (host construction in
shafft_amd/csrc/host/fastdiv.hpp) against exact division."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SRC = r"""
#include "shafft_amd/csrc/host/fastdiv.hpp"
#include <cstdio>
#include <random>
int main() {
  std::mt19937_64 g(7);
  long bad = 0;
  for (uint32_t d = 1; d < 5000; ++d) {
    sa_fastdiv_t f = sa::make_fastdiv(d);
    for (uint32_t x = 0; x < 20000; ++x) bad += sa::fastdiv_host(x, f) != x / d;
    for (int i = 0; i < 2000; ++i) { uint32_t x = g() & 0x7fffffffu; bad += sa::fastdiv_host(x, f) != x / d; }
    bad += sa::fastdiv_host(0x7fffffffu, f) != 0x7fffffffu / d;
  }
  for (int i = 0; i < 200000; ++i) {
    uint32_t d = 1 + (g() % 0x7ffffffeu), x = g() & 0x7fffffffu;
    sa_fastdiv_t f = sa::make_fastdiv(d);
    bad += sa::fastdiv_host(x, f) != x / d;
  }
  std::printf("%ld\n", bad);
  return bad != 0;
}
"""


def test_fastdiv_exact(tmp_path):
    src = tmp_path / "fd.cpp"
    src.write_text(SRC)
    exe = tmp_path / "fd"
    subprocess.run(["g++", "-O2", "-std=c++17", f"-I{ROOT}", str(src), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "0", r.stdout
