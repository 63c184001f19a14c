// Host-side check of the split-FP16 GEMM's power-of-two scale helpers
// (rapport_amd/csrc/gemm_mixed.hpp): prints pow2f(e) for e in [-160, 140] and
// scale_exp(x) for a list of probe values, one per line, for tests/test_host_cpu.py.
#include <cmath>
#include <cstdio>
#include <limits>

#include "gemm_mixed.hpp"

int main() {
  for (int e = -160; e <= 140; ++e) std::printf("p %d %a\n", e, (double)fpgemm16::pow2f(e));
  const float probes[] = {0.0f, -1.0f, 1.0f, 4095.0f, 8192.0f, 1e-30f, 1e-40f, 1.4e-45f,
                          3.0e38f, std::numeric_limits<float>::infinity(),
                          std::numeric_limits<float>::quiet_NaN(), 0.75f, 2.5e-20f, 6.0e25f};
  for (float x : probes) std::printf("s %a %d\n", (double)x, fpgemm16::scale_exp(x));
  return 0;
}
