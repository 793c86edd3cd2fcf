"""The multiply-shift division the step kernel uses for cell coordinates
(openlb_amd/csrc/fastdiv.hpp). It is compiled for the host with g++ and checked
against integer division over every divisor up to 4096, a spread of large
divisors, and the numerator edge cases of each."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r"""
#include <cstdio>
#include <cstdint>
#include <random>
#include "fastdiv.hpp"
int main() {
  std::mt19937_64 rng(7);
  long bad = 0, checked = 0;
  auto check = [&](uint32_t d, uint32_t n) {
    lbmh::FastDiv f(d);
    ++checked;
    if (f.div(n) != n / d) { if (bad < 5) std::printf("d=%u n=%u got %u\n", d, n, f.div(n)); ++bad; }
  };
  const uint32_t nmax = 0x7fffffffu;
  for (uint32_t d = 1; d <= 4096; ++d) {
    for (uint32_t n : {0u, 1u, d - 1, d, d + 1, 2 * d - 1, 2 * d, nmax, nmax - 1, nmax - d, nmax / d * d, nmax / d * d - 1})
      if (n <= nmax) check(d, n);
    for (int k = 0; k < 64; ++k) check(d, uint32_t(rng() % (uint64_t(nmax) + 1)));
  }
  for (int k = 0; k < 20000; ++k) {
    uint32_t d = uint32_t(rng() % nmax) + 1;
    for (uint32_t n : {0u, d - 1, d, nmax, nmax / d * d, nmax / d * d - 1}) check(d, n);
    for (int j = 0; j < 16; ++j) check(d, uint32_t(rng() % (uint64_t(nmax) + 1)));
  }
  for (uint32_t s = 0; s < 31; ++s) {
    const uint32_t p = 1u << s;
    for (uint32_t d : {p, p + 1, p - 1 ? p - 1 : 1u})
      for (uint32_t n : {0u, d - 1, d, nmax, nmax - 1}) check(d, n);
  }
  std::printf("checked %ld bad %ld\n", checked, bad);
  return bad != 0;
}
"""


def test_fastdiv_matches_integer_division(tmp_path):
    src = tmp_path / "fastdiv_check.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "fastdiv_check"
    subprocess.run(["g++", "-O2", "-std=c++17", f"-I{os.path.join(ROOT, 'openlb_amd', 'csrc')}",
                    str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "bad 0" in out.stdout
