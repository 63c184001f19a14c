// Host-side check of the exact float transform's compact twiddle-table layout
// (rapport_amd/csrc/fft.hpp): prints, for each N and E, the table size and the offset of
// every twiddled pass, for tests/test_host_cpu.py to compare with the host-built table.
#include <cstdio>

#include "fft.hpp"

template <int N, int E, int Ns>
void pass_offsets() {
  if constexpr (Ns < N) {
    constexpr int R = N / Ns >= E ? E : N / Ns;
    if constexpr (Ns > 1) std::printf("o %d %d %d %d %d\n", N, E, Ns, R, fpfft::exact_table_offset<N, E, Ns>());
    pass_offsets<N, E, Ns * R>();
  }
}

template <int N>
void one() {
  std::printf("t %d %d %d %d\n", N, fpfft::exact_pass_entries(N, 8, 1), fpfft::exact_pass_entries(N, 4, 1),
              fpfft::exact_pass_entries(N, 16, 1));
  pass_offsets<N, 8, 1>();
  pass_offsets<N, 4, 1>();
  pass_offsets<N, 16, 1>();
}

int main() {
  one<512>();
  one<1024>();
  one<2048>();
  one<4096>();
  return 0;
}
