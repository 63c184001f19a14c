// Hankel transform as a real x complex GEMM on FP64 MFMA (gfx950).
//
// C (M x Ncols) = alpha * diag(rowscale) * Y (M x K) @ B (K x Ncols), where B
// and C are complex rows viewed as real (Ncols = 2 Nt, columns alternate re/im)
// and Y is real, so one real GEMM does the whole complex product. Optionally
// the epilogue multiplies each complex output (m, f) by phase[m, f].
//
// Tiling: 256 threads = 4 waves per 128 x 128 tile of C; each wave owns a
// 64 x 64 quadrant as 4 x 4 blocks of v_mfma_f64_16x16x4f64. K advances in
// slabs of 16 staged through LDS with double buffering (global -> registers
// for slab s+1 while MFMAs run on slab s). A is read from Y^T (K x M) so both
// LDS tiles are k-major and every MFMA operand read is a contiguous 16-lane
// row. Block indices are remapped so that each XCD works on a contiguous run
// of column panels (all row tiles of a panel land on one XCD's L2).
#pragma once
#include "common.hpp"

namespace fpgemm {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int BM = 128, BN = 128, BK = 16, NTHR = 256;
constexpr int APAD = 4;  // LDS row padding (doubles)

__global__ __launch_bounds__(NTHR, 2) void hankel_gemm_f64(const double* __restrict__ Yt,
                                                         const double* __restrict__ B,
                                                         double* __restrict__ C, int M, int K,
                                                         int Ncols, double alpha,
                                                         const double* __restrict__ rowscale,
                                                         const double* __restrict__ phase) {
  __shared__ __attribute__((aligned(16))) double As[2][BK][BM + APAD];
  __shared__ __attribute__((aligned(16))) double Bs[2][BK][BN + APAD];

  const int tilesM = (M + BM - 1) / BM;
  const int tilesN = (Ncols + BN - 1) / BN;
  const int nb = tilesM * tilesN;
  // XCD-aware remap: hardware places block id on XCD (id % 8).
  int id = blockIdx.x;
  int lin = id;
  if (nb % 8 == 0) lin = (id % 8) * (nb / 8) + id / 8;
  const int tm = lin % tilesM, tn = lin / tilesM;
  const int m0 = tm * BM, n0 = tn * BN;

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;

  // global->LDS staging: each thread moves 4 double pairs of A (row ldk of the Y^T slab)
  // and 4 of B per slab, pair i at column ldc + 32 i: the 16 threads of a row cover 256
  // contiguous bytes per pair, so both the global loads and the 16-byte LDS stores are
  // contiguous across lanes (no LDS bank conflicts); full tiles use 16-byte accesses.
  const int ldk = tid >> 4;
  const int ldc = (tid & 15) * 2;
  const bool full = (m0 + BM <= M) && (n0 + BN <= Ncols) && (K % BK == 0);
  double2 ra[4], rb[4];

  auto gload = [&](int k0) {
    const int k = k0 + ldk;
    if (full) {
      const double* pa = Yt + (size_t)k * M + m0 + ldc;
      const double* pb = B + (size_t)k * Ncols + n0 + ldc;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        ra[i] = *reinterpret_cast<const double2*>(pa + 32 * i);
        rb[i] = *reinterpret_cast<const double2*>(pb + 32 * i);
      }
    } else {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int m = m0 + ldc + 32 * i, n = n0 + ldc + 32 * i;
        ra[i].x = (k < K && m < M) ? Yt[(size_t)k * M + m] : 0.0;
        ra[i].y = (k < K && m + 1 < M) ? Yt[(size_t)k * M + m + 1] : 0.0;
        rb[i].x = (k < K && n < Ncols) ? B[(size_t)k * Ncols + n] : 0.0;
        rb[i].y = (k < K && n + 1 < Ncols) ? B[(size_t)k * Ncols + n + 1] : 0.0;
      }
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      *reinterpret_cast<double2*>(&As[buf][ldk][ldc + 32 * i]) = ra[i];
      *reinterpret_cast<double2*>(&Bs[buf][ldk][ldc + 32 * i]) = rb[i];
    }
  };

  d4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};

  const int nslab = (K + BK - 1) / BK;
  gload(0);
  sstore(0);
  __syncthreads();
  const int li = lane & 15, lk = lane >> 4;

  // Phase values: every lane loads the double at its own output (m, n) -- re for even n,
  // im for odd n -- and takes the other part of the complex phase from lane n ^ 1 by a
  // shuffle. Row block a = 0 is loaded one K slab before the end, so it streams in under
  // the last slab's MFMAs (the 512 tiles reach the epilogue together, with no work left to
  // overlap a 64 MiB table read); in the epilogue block a + 1 is loaded before block a is
  // written, so at most two blocks of phase values are live beside the accumulators.
  double php[4][4][4];
  auto pload = [&](int a) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int m = m0 + wm + a * 16 + lk + 4 * v;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int n = n0 + wn + b * 16 + li;
        php[a][v][b] = (m < M && n < Ncols) ? phase[(size_t)m * Ncols + n] : 0.0;
      }
    }
  };
  const int pf_slab = nslab >= 2 ? nslab - 2 : -1;   // block 0 issued in that iteration

  for (int s = 0; s < nslab; ++s) {
    const int cur = s & 1;
    if (s + 1 < nslab) gload((s + 1) * BK);
    if (phase && s == pf_slab) pload(0);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      double av[4], bv[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) av[a] = As[cur][kk + lk][wm + a * 16 + li];
#pragma unroll
      for (int b = 0; b < 4; ++b) bv[b] = Bs[cur][kk + lk][wn + b * 16 + li];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
    if (s + 1 < nslab) {
      sstore(cur ^ 1);
      __syncthreads();
    }
  }

  if (phase && pf_slab < 0) pload(0);

  // epilogue: for v_mfma_f64_16x16x4f64 lane l, register v holds D[l/16 + 4 v][l % 16]
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    if (phase && a + 1 < 4) pload(a + 1);   // lands while block a is written
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int m = m0 + wm + a * 16 + lk + 4 * v;
      const double rs = (rowscale && m < M) ? alpha * rowscale[m] : alpha;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int n = n0 + wn + b * 16 + li;
        double val = acc[a][b][v] * rs;
        if (phase) {
          // even lane holds re, odd lane holds im of the same complex element (m, n >> 1)
          const double other = __shfl_xor(val, 1, 64);
          const double own_ph = php[a][v][b];
          const double other_ph = __shfl_xor(own_ph, 1, 64);
          const double pr = (n & 1) ? other_ph : own_ph;
          const double pi = (n & 1) ? own_ph : other_ph;
          val = (n & 1) ? (other * pi + val * pr) : (val * pr - other * pi);
        }
        if (m < M && n < Ncols) C[(size_t)m * Ncols + n] = val;
      }
    }
  }
}

// Column band of the FP64 GEMM (see the mixed GEMM's ColBand): virtual column v is real
// column v + (v >= va ? shift : 0).
struct Band64 {
  int va, shift;
  FP_DEV int map(int v) const { return v >= va ? v + shift : v; }
};

// The FP64 GEMM restricted to a column band: tiles of 128 rows x TBK 16-column blocks,
// `slots` tiles per row panel over the band's virtual columns, two blocks per CU. The four
// waves each take 32 rows and all TBK blocks (equal work per SIMD whatever TBK is). Same
// operand staging, K order, phase epilogue and per-element arithmetic as hankel_gemm_f64,
// so the computed columns are bitwise its values. Needs M % 128 == 0. Afterwards each
// block writes zeros to its share of the real columns [zero_lo, zero_hi).
template <int TBK>
__global__ __launch_bounds__(NTHR, 2) void hankel_gemm_f64_band(
    const double* __restrict__ Yt, const double* __restrict__ B, double* __restrict__ C, int M,
    int K, int Ncols, double alpha, const double* __restrict__ rowscale,
    const double* __restrict__ phase, int slots, Band64 band, int zero_lo, int zero_hi) {
  __shared__ __attribute__((aligned(16))) double As[2][BK][BM + APAD];
  __shared__ __attribute__((aligned(16))) double Bs[2][BK][BN + APAD];
  constexpr int FA = 2, NCB = TBK * 16;
  const int tilesM = M / BM;
  const int nb = tilesM * slots;
  int lin = blockIdx.x;
  if (nb % 8 == 0) lin = (lin % 8) * (nb / 8) + lin / 8;
  const int tm = lin % tilesM, sl = lin / tilesM;
  const int m0 = tm * BM, n0 = sl * NCB;          // n0: virtual column

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave * 32;

  const int ldk = tid >> 4;
  const int ldc = (tid & 15) * 2;
  int bcol[4];                                     // real column of each staged pair
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int c = ldc + 32 * i;
    bcol[i] = band.map(n0 + (c < NCB ? c : NCB - 2));
  }
  double2 ra[4], rb[4];
  auto gload = [&](int k0) {
    const int k = k0 + ldk;
    const double* pa = Yt + (size_t)k * M + m0 + ldc;
    const double* pb = B + (size_t)k * Ncols;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (k < K) {
        ra[i] = *reinterpret_cast<const double2*>(pa + 32 * i);
        rb[i] = *reinterpret_cast<const double2*>(pb + bcol[i]);
      } else {
        ra[i] = double2{0.0, 0.0};
        rb[i] = double2{0.0, 0.0};
      }
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      *reinterpret_cast<double2*>(&As[buf][ldk][ldc + 32 * i]) = ra[i];
      *reinterpret_cast<double2*>(&Bs[buf][ldk][ldc + 32 * i]) = rb[i];
    }
  };

  d4 acc[FA][TBK];
#pragma unroll
  for (int a = 0; a < FA; ++a)
#pragma unroll
    for (int b = 0; b < TBK; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};

  const int nslab = (K + BK - 1) / BK;
  gload(0);
  sstore(0);
  __syncthreads();
  const int li = lane & 15, lk = lane >> 4;

  for (int s = 0; s < nslab; ++s) {
    const int cur = s & 1;
    if (s + 1 < nslab) gload((s + 1) * BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      double av[FA], bv[TBK];
#pragma unroll
      for (int a = 0; a < FA; ++a) av[a] = As[cur][kk + lk][wm + a * 16 + li];
#pragma unroll
      for (int b = 0; b < TBK; ++b) bv[b] = Bs[cur][kk + lk][b * 16 + li];
#pragma unroll
      for (int a = 0; a < FA; ++a)
#pragma unroll
        for (int b = 0; b < TBK; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
    if (s + 1 < nslab) {
      sstore(cur ^ 1);
      __syncthreads();
    }
  }

  // epilogue: the phase values of one 16-row block at a time (wider tiles leave no
  // registers to prefetch them under the last slab, as hankel_gemm_f64 does)
#pragma unroll
  for (int a = 0; a < FA; ++a) {
    double php[4][TBK];
    if (phase) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int m = m0 + wm + a * 16 + lk + 4 * v;
#pragma unroll
        for (int b = 0; b < TBK; ++b)
          php[v][b] = phase[(size_t)m * Ncols + band.map(n0 + b * 16 + li)];
      }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int m = m0 + wm + a * 16 + lk + 4 * v;
      const double rs = rowscale ? alpha * rowscale[m] : alpha;
#pragma unroll
      for (int b = 0; b < TBK; ++b) {
        const int n = band.map(n0 + b * 16 + li);
        double val = acc[a][b][v] * rs;
        if (phase) {
          const double other = __shfl_xor(val, 1, 64);
          const double own_ph = php[v][b];
          const double other_ph = __shfl_xor(own_ph, 1, 64);
          const double pr = (n & 1) ? other_ph : own_ph;
          const double pi = (n & 1) ? own_ph : other_ph;
          val = (n & 1) ? (other * pi + val * pr) : (val * pr - other * pi);
        }
        C[(size_t)m * Ncols + n] = val;
      }
    }
  }
  if (zero_hi > zero_lo) {
    const int w2 = (zero_hi - zero_lo) / 2;
    const long long total = (long long)M * w2;
    const long long chunk = (total + gridDim.x - 1) / gridDim.x;
    const long long end = min(total, (long long)(blockIdx.x + 1) * chunk);
    for (long long i = (long long)blockIdx.x * chunk + threadIdx.x; i < end; i += NTHR) {
      const int r = (int)(i / w2), c2 = (int)(i % w2);
      *reinterpret_cast<double2*>(C + (size_t)r * Ncols + zero_lo + 2 * c2) = double2{0.0, 0.0};
    }
  }
}

}  // namespace fpgemm
