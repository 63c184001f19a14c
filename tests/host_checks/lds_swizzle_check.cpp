// Host-side check of the row transforms' LDS swizzle (rapport_amd/csrc/fft.hpp, lds_pad):
// for every row shape the library builds (N = 512 .. 4096 samples, 4, 8 or 16 samples per
// thread) it checks that lds_pad is a bijection on [0, N), that it is linear over XOR, that
// the per-butterfly write addresses (base XOR constant) and the stride-layout read addresses
// (base XOR mask + offset) the transforms use equal lds_pad of the element index, and it
// prints the worst bank-conflict degree of the Stockham writes and stride-layout reads of
// every wave under a 32-bank x 4-byte LDS model (16 lanes per cycle for 8-byte elements,
// 8 for 16-byte ones), one line per (N, threads, element size), for tests/test_host_cpu.py.
#include <algorithm>
#include <cstdio>
#include <set>
#include <vector>

#include "fft.hpp"

using fpfft::lds_mask;
using fpfft::lds_pad;

static int degree(const std::vector<int>& elems, int es) {
  constexpr int kBanks = 32;
  const int lpc = kBanks * 4 / es;
  int worst = 0;
  for (size_t h = 0; h < elems.size(); h += lpc) {
    int count[kBanks] = {0};
    std::set<int> words;
    for (size_t l = h; l < h + lpc && l < elems.size(); ++l)
      for (int w = 0; w < es / 4; ++w) {
        const int word = elems[l] * (es / 4) + w;
        if (words.insert(word).second) ++count[word % kBanks];
      }
    worst = std::max(worst, *std::max_element(count, count + kBanks));
  }
  return worst;
}

int main() {
  int fails = 0;
  for (int N = 512; N <= 4096; N *= 2) {
    std::vector<int> seen(N, 0);
    for (int i = 0; i < N; ++i) {
      const int p = lds_pad(i);
      if (p < 0 || p >= N || seen[p]++) ++fails;
    }
    for (int a = 0; a < N; a += 7)
      for (int b = 0; b < N; ++b)
        if (lds_pad(a ^ b) != (lds_pad(a) ^ lds_pad(b))) ++fails;
    for (int E = 4; E <= 16; E *= 2) {
      const int TPB = N / E;
      if (TPB < 64) continue;
      for (int es = 8; es <= 16; es *= 2) {
        int wdeg = 0, rdeg = 0;
        for (int Ns = 1;;) {
          const int R = N / Ns >= E ? E : N / Ns, Q = E / R;
          if (Ns * R == N) break;
          for (int w0 = 0; w0 < TPB; w0 += 64) {
            for (int q = 0; q < Q; ++q)
              for (int r = 0; r < R; ++r) {
                std::vector<int> el;
                for (int l = 0; l < 64; ++l) {
                  const int j = w0 + l + q * TPB, k = j & (Ns - 1);
                  const int idxD = (j / Ns) * Ns * R + k;
                  if ((lds_pad(idxD) ^ lds_pad(r * Ns)) != lds_pad(idxD + r * Ns)) ++fails;
                  el.push_back(lds_pad(idxD + r * Ns));
                }
                wdeg = std::max(wdeg, degree(el, es));
              }
            for (int s = 0; s < E; ++s) {
              std::vector<int> el;
              for (int l = 0; l < 64; ++l) {
                const int tid = w0 + l;
                if (TPB % 256 == 0 &&
                    ((lds_pad(tid) ^ lds_mask(s * TPB)) + s * TPB) != lds_pad(tid + s * TPB))
                  ++fails;
                el.push_back(lds_pad(tid + s * TPB));
              }
              rdeg = std::max(rdeg, degree(el, es));
            }
          }
          Ns *= R;
        }
        std::printf("c %d %d %d %d %d\n", N, TPB, es, wdeg, rdeg);
      }
    }
  }
  std::printf("f %d\n", fails);
  return 0;
}
