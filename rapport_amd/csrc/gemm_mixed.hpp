// Hankel GEMM in mixed precision: split-FP16 operands on MFMA, FP32 accumulation.
//
// C (M x Ncols, float) = alpha * diag(rowscale) * Y @ B (optionally x phase[m, f]),
// with B a complex64 field viewed as M x Ncols floats. Both operands are split
// into FP16 hi + lo parts, Y = (Yh + Yl)/sY and B = (Bh + Bl)/sB, and the
// product is formed as Yh Bh + Yh Bl + Yl Bh (the Yl Bl term is below FP32
// resolution), i.e. three v_mfma_f32_16x16x32_f16 per 16x16x32 block. The
// scales are powers of two: sY is fixed when Y is split on the host, sB comes
// from max|B|, which the kernel that produced B left in a device word (float
// bits, atomicMax), so that |B sB| < 2^13 and no FP16 overflow can occur.
//
// Y is stored pre-split in MFMA fragment order, slab-major:
//   Yh[ks][mb][kg][r][8], ks = K/32 slab, mb = M/16 row block, kg = k/8 group,
//   r = row in block; element j of lane (kg*16 + r) is Y[16 mb + r][32 ks + 8 kg + j].
// B is split on the fly while staging into LDS in the same fragment order.
//
// Tile 128 x 128 per 256-thread block (4 waves x 64 x 64), K slab 32, LDS
// double-buffered (64 KiB), two blocks per CU. The epilogue also reduces
// max|C| into a device word for the next GEMM in the chain.
#pragma once
#include "common.hpp"

namespace fpgemm16 {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int BM = 128, BN = 128, BK = 32, NTHR = 256;
constexpr int SLAB_HALVES = 8 * 4 * 16 * 8;   // 128 rows (or cols) x 32 k

FP_DEV float pow2f(int e) { return __int_as_float((e + 127) << 23); }

// Power-of-two scale s with max * s < 2^13 (max >= 0); 1 when max is 0 / not finite.
FP_DEV int scale_exp(float mx) {
  if (!(mx > 0.0f) || !fp_isfinite(mx)) return 0;
  int e;
  frexpf(mx, &e);           // mx < 2^e
  int se = 13 - e;
  if (se > 100) se = 100;
  if (se < -100) se = -100;
  return se;
}

__global__ __launch_bounds__(NTHR, 2) void hankel_gemm_f16x3(
    const h8* __restrict__ Yh, const h8* __restrict__ Yl, int y_scale_exp,
    const float* __restrict__ B, float* __restrict__ C, int M, int K, int Ncols, float alpha,
    const float* __restrict__ rowscale, const float* __restrict__ phase,
    const unsigned* __restrict__ bmax_in, unsigned* __restrict__ cmax_out,
    unsigned* __restrict__ reset_slot) {
  __shared__ h8 sAh[2][SLAB_HALVES / 8], sAl[2][SLAB_HALVES / 8];
  __shared__ h8 sBh[2][SLAB_HALVES / 8], sBl[2][SLAB_HALVES / 8];
  __shared__ float sred[NTHR / 64];

  const int tid = threadIdx.x;
  if (reset_slot && blockIdx.x == 0 && tid == 0) *reset_slot = 0u;

  const int tilesM = M / BM, tilesN = Ncols / BN;
  const int nb = tilesM * tilesN;
  int lin = blockIdx.x;
  if (nb % 8 == 0) lin = (blockIdx.x % 8) * (nb / 8) + blockIdx.x / 8;   // XCD-contiguous panels
  const int tm = lin % tilesM, tn = lin / tilesM;
  const int m0 = tm * BM, n0 = tn * BN;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;

  const int sb_exp = scale_exp(__uint_as_float(*bmax_in));
  const float sB = pow2f(sb_exp);

  // staging roles
  const int bc = tid & 127;            // B column within the tile
  const int bkp = tid >> 7;            // k-group pair (groups 2 bkp, 2 bkp + 1)
  float rb[16];
  h8 ra_h[2], ra_l[2];
  const size_t a_slab_stride = (size_t)(M / 16) * 64;   // h8 entries per K slab
  const size_t a_tile_off = (size_t)(m0 / 16) * 64;

  auto gload = [&](int ks) {
    const int k0 = ks * BK;
#pragma unroll
    for (int g = 0; g < 2; ++g) {
      const int kg = 2 * bkp + g;
#pragma unroll
      for (int j = 0; j < 8; ++j) rb[g * 8 + j] = B[(size_t)(k0 + 8 * kg + j) * Ncols + n0 + bc];
    }
    const size_t base = (size_t)ks * a_slab_stride + a_tile_off;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      ra_h[i] = Yh[base + tid + i * NTHR];
      ra_l[i] = Yl[base + tid + i * NTHR];
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int g = 0; g < 2; ++g) {
      const int kg = 2 * bkp + g;
      h8 hi, lo;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const float x = rb[g * 8 + j] * sB;
        const _Float16 h = (_Float16)x;
        hi[j] = h;
        lo[j] = (_Float16)(x - (float)h);
      }
      const int idx = ((bc >> 4) * 4 + kg) * 16 + (bc & 15);
      sBh[buf][idx] = hi;
      sBl[buf][idx] = lo;
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      sAh[buf][tid + i * NTHR] = ra_h[i];
      sAl[buf][tid + i * NTHR] = ra_l[i];
    }
  };

  f4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = f4{0.f, 0.f, 0.f, 0.f};

  const int nslab = K / BK;
  gload(0);
  sstore(0);
  __syncthreads();
  const int lq = lane >> 4, lr = lane & 15;
  for (int s = 0; s < nslab; ++s) {
    const int cur = s & 1;
    if (s + 1 < nslab) gload(s + 1);
    h8 ah[4], al[4], bh[4], bl[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int idx = (((wm >> 4) + a) * 4 + lq) * 16 + lr;
      ah[a] = sAh[cur][idx];
      al[a] = sAl[cur][idx];
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int idx = (((wn >> 4) + b) * 4 + lq) * 16 + lr;
      bh[b] = sBh[cur][idx];
      bl[b] = sBl[cur][idx];
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[a], bh[b], acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[a], bl[b], acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[a], bh[b], acc[a][b], 0, 0, 0);
      }
    if (s + 1 < nslab) {
      sstore(cur ^ 1);
      __syncthreads();
    }
  }

  // epilogue: for v_mfma_f32_16x16x32_f16 lane l, register v holds D[4 (l/16) + v][l % 16]
  const float unscale = pow2f(-(y_scale_exp + sb_exp));
  float vmax = 0.0f;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int m = m0 + wm + a * 16 + 4 * lq + v;
      const float rs = rowscale ? alpha * unscale * rowscale[m] : alpha * unscale;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int n = n0 + wn + b * 16 + lr;
        float val = acc[a][b][v] * rs;
        if (phase) {
          const float other = __shfl_xor(val, 1, 64);
          const size_t pidx = ((size_t)m * (Ncols >> 1) + (n >> 1)) * 2;
          const float pr = phase[pidx], pi = phase[pidx + 1];
          val = (n & 1) ? (other * pi + val * pr) : (val * pr - other * pi);
        }
        vmax = fmaxf(vmax, fabsf(val));
        C[(size_t)m * Ncols + n] = val;
      }
    }
  }
  if (cmax_out) {
    for (int d = 32; d >= 1; d >>= 1) vmax = fmaxf(vmax, __shfl_xor(vmax, d, 64));
    if (lane == 0) sred[wave] = vmax;
    __syncthreads();
    if (tid == 0) {
      float m = sred[0];
      for (int w = 1; w < NTHR / 64; ++w) m = fmaxf(m, sred[w]);
      atomicMax(cmax_out, __float_as_uint(m));
    }
  }
}

}  // namespace fpgemm16
