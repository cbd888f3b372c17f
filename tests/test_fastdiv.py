"""The solver's division by invariant integers (csrc/fastdiv.h) against
plain unsigned division, compiled for the host: exhaustive over small
dividends and random over the 32-bit range, for the divisor shapes the
solver uses (nodes per layer, nodes per expanded layer) and edge cases."""
import os
import subprocess
import tempfile

HDR = os.path.join(os.path.dirname(__file__), "..", "product-graph-cycles_amd", "csrc")

SRC = r"""
#include "fastdiv.h"
#include <cstdio>
#include <random>
int main() {
  std::mt19937_64 r(7);
  long bad = 0;
  const uint32_t ds[] = {1u, 2u, 3u, 5u, 7u, 160u, 162u, 180u, 642u, 720u, 1284u, 2562u, 2880u, 7686u, 11520u,
                         46080u, 92160u, 138240u, 65535u, 65536u, 65537u, 0x7fffffffu, 0x80000000u, 0xffffffffu};
  for (uint32_t d : ds) {
    pgc::FastDiv f;
    f.init(d);
    for (uint32_t x = 0; x < 200000; ++x) bad += f.div(x) != x / d;
    for (int i = 0; i < 200000; ++i) {
      const uint32_t x = (uint32_t)r();
      bad += f.div(x) != x / d;
    }
  }
  for (int k = 0; k < 5000; ++k) {
    uint32_t d = (uint32_t)(r() >> (r() % 64));
    if (!d) d = 1;
    pgc::FastDiv f;
    f.init(d);
    for (int i = 0; i < 100; ++i) {
      const uint32_t x = (uint32_t)r();
      bad += f.div(x) != x / d;
    }
  }
  std::printf("%ld\n", bad);
  return bad != 0;
}
"""


def test_fastdiv_matches_division():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "t.cpp")
        exe = os.path.join(d, "t")
        with open(src, "w") as f:
            f.write(SRC)
        subprocess.run(["g++", "-O2", "-std=c++17", "-I", HDR, src, "-o", exe], check=True)
        r = subprocess.run([exe], capture_output=True, text=True)
        assert r.returncode == 0 and r.stdout.strip() == "0", r.stdout
