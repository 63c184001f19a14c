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
// Tile BM x BN per block of WM x WN waves, each (BM/WM) x (BN/WN): 2 x 2 waves with
// BM, BN in {64, 128} (or 128 x 256), two blocks per CU; or 4 x 2 waves of 64 x 64 for a
// 256 x 128 tile, one block per CU (each B column panel is then split by half as many row
// tiles). K slab 32, LDS double-buffered. The epilogue also reduces max|C| into a device
// word for the next GEMM in the chain.
#pragma once
#include <type_traits>
#include <utility>

#include "common.hpp"

namespace fpgemm16 {

// f(std::integral_constant<int, i>) for i = 0 .. N-1, unrolled at compile time (register
// arrays indexed by i stay in registers).
template <typename F, int... I>
FP_DEV void static_for_seq(F&& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
FP_DEV void static_for(F&& f) {
  static_for_seq(f, std::make_integer_sequence<int, N>{});
}

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int BK = 32, NTHR = 256;

// 2^e as a float for any int e: a normal number for -126 <= e <= 127, subnormal down to
// 2^-149, 0 below that and +inf above 127 (host and device).
__host__ __device__ inline float pow2f(int e) {
  if (e > 127) return __builtin_bit_cast(float, 0x7f800000u);
  if (e >= -126) return __builtin_bit_cast(float, (unsigned)(e + 127) << 23);
  if (e >= -149) return __builtin_bit_cast(float, 1u << (e + 149));
  return 0.0f;
}

// Power-of-two scale s with max * s < 2^13 (max >= 0); 1 when max is 0 / not finite.
__host__ __device__ inline int scale_exp(float mx) {
  if (!(mx > 0.0f) || !__builtin_isfinite(mx)) return 0;
  int e;
  frexpf(mx, &e);           // mx < 2^e
  int se = 13 - e;
  if (se > 100) se = 100;
  if (se < -100) se = -100;
  return se;
}

// B's power-of-two scale from the max|B| word. The word is read with a vector load from
// an index the compiler cannot prove uniform: a scalar load would be waited for together
// with the kernel arguments, before the first operand load is issued, which puts one more
// memory round trip in front of every tile. Callers issue their first operand loads first.
FP_DEV void load_scale(const unsigned* __restrict__ bmax_in, int& sb_exp, float& sB) {
  int z = 0;
  asm volatile("" : "+v"(z));
  sb_exp = scale_exp(__uint_as_float(bmax_in[z]));
  sB = pow2f(sb_exp);
}

// One K slab of a BM-row (BN-column) tile is BM * 4 h8 fragments.
template <int BM, int BN, int WM = 2, int WN = 2> struct Smem {
  h8 Ah[2][BM * 4], Al[2][BM * 4];
  h8 Bh[2][BN * 4], Bl[2][BN * 4];
  float red[WM * WN];
};

// Column band of a tile: the tile's columns are a run of a "virtual" column space that
// skips the frequencies where the spectral mask is exactly 0 (real columns [va, va + shift)
// of C and B are left out: virtual column v is real column v + (v >= va ? shift : 0)).
struct ColBand {
  int va, shift;
  FP_DEV int map(int v) const { return v >= va ? v + shift : v; }
};
constexpr ColBand kNoBand = {1 << 30, 0};

// One BM x BN output tile (tm, tn) of C = alpha diag(rowscale) Y B (x phase). With
// TBK < BN / 16 the tile is TBK 16-column blocks wide (the rest of the B panel is staged
// but not multiplied or stored); the two wave columns then take ceil(TBK / 2) and
// floor(TBK / 2) blocks, and wave w has wave row w % WM and wave column w / WM, so each
// SIMD (w % 4) holds one wave of each column and the SIMDs stay balanced. The tile's first
// virtual column is n0v.
template <int BM, int BN, int WM = 2, int WN = 2, int NPF = 0, int TBK = BN / 16>
FP_DEV void gemm_tile(Smem<BM, BN, WM, WN>& sm, int tm, int tn, const h8* __restrict__ Yh,
                      const h8* __restrict__ Yl, int y_scale_exp, const float* __restrict__ B,
                      float* __restrict__ C, int M, int K, int Ncols, float alpha,
                      const float* __restrict__ rowscale, const float* __restrict__ phase,
                      const unsigned* __restrict__ bmax_in, unsigned* __restrict__ cmax_out,
                      ColBand band = kNoBand, int n0v = -1) {
  const int tid = threadIdx.x;
  const int m0 = tm * BM, n0 = n0v >= 0 ? n0v : tn * BN;
  const int lane = tid & 63, wave = tid >> 6;
  constexpr int NT = WM * WN * 64;
  constexpr int FA = BM / WM / 16, FB = BN / WN / 16;   // 16-row / 16-col blocks per wave
  constexpr int CPT = NT / BN;                     // threads per B column
  constexpr int GPT = 4 / CPT;                     // k groups of 8 staged per thread
  constexpr int APT = BM * 4 / NT;                 // A fragments staged per thread
  constexpr bool NARROW = TBK < BN / 16;
  static_assert(GPT >= 1 && APT >= 1, "tile too small for the block");
  static_assert(!NARROW || (WN == 2 && (TBK + 1) / 2 <= FB), "narrow tiles: two wave columns");
  const int wm = NARROW ? (wave % WM) * (BM / WM) : (wave / WN) * (BM / WM);
  // narrow tiles: wave column 1 starts right after wave column 0's ceil(TBK / 2) blocks
  const int wn = NARROW ? (wave / WM) * ((TBK + 1) / 2) * 16 : (wave % WN) * (BN / WN);
  const bool wcol0 = wn == 0;

  // B's scale is read after the first operand loads are issued (load_scale below)
  int sb_exp = 0;
  float sB = 1.0f;

  // staging roles
  const int bc = tid % BN;             // B column within the tile
  const int bkp = tid / BN;            // k groups GPT bkp .. GPT bkp + GPT - 1
  const size_t a_slab_stride = (size_t)(M / 16) * 64;   // h8 entries per K slab
  const size_t a_tile_off = (size_t)(m0 / 16) * 64;

  // Two register staging sets: the global loads of slab s + 2 are issued while
  // slab s is multiplied, so each load has two slabs of MFMA work to land.
  struct Stage {
    float rb[8 * GPT];
    h8 ra_h[APT], ra_l[APT];
  };
  Stage R0, R1;

  // 32-bit byte offsets from the uniform base pointers (every operand here is below
  // 4 GiB): loads take the scalar-base + 32-bit vector-offset form, no 64-bit address math.
  const char* Bc = reinterpret_cast<const char*>(B);
  const char* Yhc = reinterpret_cast<const char*>(Yh);
  const char* Ylc = reinterpret_cast<const char*>(Yl);
  const unsigned bcol4 = 4u * (unsigned)band.map(n0 + (bc < TBK * 16 ? bc : TBK * 16 - 1));
  const unsigned row4 = 4u * (unsigned)Ncols;
  auto gload = [&](int ks, Stage& R) {
    const unsigned k0 = (unsigned)(ks * BK);
#pragma unroll
    for (int g = 0; g < GPT; ++g) {
      const unsigned kg = (unsigned)(GPT * bkp + g);
      const unsigned r0 = (k0 + 8u * kg) * row4 + bcol4;
#pragma unroll
      for (int j = 0; j < 8; ++j)
        R.rb[g * 8 + j] = *reinterpret_cast<const float*>(Bc + (r0 + (unsigned)j * row4));
    }
    const unsigned base = 16u * ((unsigned)(ks * a_slab_stride + a_tile_off) + (unsigned)tid);
#pragma unroll
    for (int i = 0; i < APT; ++i) {
      R.ra_h[i] = *reinterpret_cast<const h8*>(Yhc + (base + 16u * (unsigned)(i * NT)));
      R.ra_l[i] = *reinterpret_cast<const h8*>(Ylc + (base + 16u * (unsigned)(i * NT)));
    }
  };
  auto sstore = [&](int buf, const Stage& R) {
#pragma unroll
    for (int g = 0; g < GPT; ++g) {
      const int kg = GPT * bkp + g;
      h8 hi, lo;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const float x = R.rb[g * 8 + j] * sB;
        const _Float16 h = (_Float16)x;
        hi[j] = h;
        lo[j] = (_Float16)(x - (float)h);
      }
      const int idx = ((bc >> 4) * 4 + kg) * 16 + (bc & 15);
      sm.Bh[buf][idx] = hi;
      sm.Bl[buf][idx] = lo;
    }
#pragma unroll
    for (int i = 0; i < APT; ++i) {
      sm.Ah[buf][tid + i * NT] = R.ra_h[i];
      sm.Al[buf][tid + i * NT] = R.ra_l[i];
    }
  };

  f4 acc[FA][FB];
#pragma unroll
  for (int a = 0; a < FA; ++a)
#pragma unroll
    for (int b = 0; b < FB; ++b) acc[a][b] = f4{0.f, 0.f, 0.f, 0.f};

  const int lq = lane >> 4, lr = lane & 15;
  // MFMA phase of a slab at raised wave priority, with the scheduler told to interleave the
  // B-fragment LDS reads of the next column block between the current block's MFMAs (both
  // measured: 10 % faster at 1024 x 4096, results bitwise unchanged). NB: column blocks
  // this wave multiplies (FB, or its share of a narrow tile).
  auto mma_n = [&](int cur, auto nb_c) {
    constexpr int NB = decltype(nb_c)::value;
    __builtin_amdgcn_s_setprio(1);
    h8 ah[FA], al[FA];
#pragma unroll
    for (int a = 0; a < FA; ++a) {
      const int idx = (((wm >> 4) + a) * 4 + lq) * 16 + lr;
      ah[a] = sm.Ah[cur][idx];
      al[a] = sm.Al[cur][idx];
    }
    // B fragments per column block: keeps the live registers at A + one B pair
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int idx = (((wn >> 4) + b) * 4 + lq) * 16 + lr;
      const h8 bh = sm.Bh[cur][idx];
      const h8 bl = sm.Bl[cur][idx];
#pragma unroll
      for (int a = 0; a < FA; ++a) {
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[a], bh, acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[a], bl, acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[a], bh, acc[a][b], 0, 0, 0);
      }
    }
    // groups: the A fragments and the first B pair (DS reads), then per column block
    // half of its 3 FA MFMAs, the next block's two B reads, the other half
    __builtin_amdgcn_sched_group_barrier(0x100, 2 * FA + 2, 0);
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      __builtin_amdgcn_sched_group_barrier(0x008, 3 * FA / 2, 0);
      if (b + 1 < NB) __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);
      __builtin_amdgcn_sched_group_barrier(0x008, 3 * FA / 2, 0);
    }
    __builtin_amdgcn_s_setprio(0);
  };
  auto mma = [&](int cur) {
    if constexpr (NARROW) {
      if (wcol0) mma_n(cur, std::integral_constant<int, (TBK + 1) / 2>{});
      else mma_n(cur, std::integral_constant<int, TBK / 2>{});
    } else {
      mma_n(cur, std::integral_constant<int, FB>{});
    }
  };
  const int fbw = NARROW ? (wcol0 ? (TBK + 1) / 2 : TBK / 2) : FB;   // blocks this wave stores

  // Phase prefetch (NPF > 0): every lane needs, for each of its outputs (m, n), the phase
  // of the complex element (n & ~1, n | 1) of row m. It loads the one float at its own
  // column n (re for even n, im for odd n) and takes the other from its neighbour lane
  // (n ^ 1) by a shuffle in the epilogue. The loads of 16-row blocks a < NPF go out right
  // after the operand loads of loop step pf0 + 2a, so that part of the phase table streams
  // in while the last slabs are multiplied; the loads of the other blocks go out at the
  // start of the epilogue (the staging registers are free by then) and land while the
  // prefetched blocks are written. NPF = 0 reads each phase pair in the epilogue loop.
  const char* Pc = reinterpret_cast<const char*>(phase);
  const unsigned ncol4 = 4u * (unsigned)Ncols;
  float php[FA][4][FB];   // unused (and removed) when NPF == 0
  auto pload = [&](auto a_c) {
    constexpr int a = decltype(a_c)::value;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const unsigned rowoff = (unsigned)(m0 + wm + a * 16 + 4 * lq + v) * ncol4;
#pragma unroll
      for (int b = 0; b < FB; ++b)
        php[a][v][b] = *reinterpret_cast<const float*>(
            Pc + rowoff + 4u * (unsigned)band.map(n0 + wn + (b < fbw ? b : 0) * 16 + lr));
    }
  };
  // in_loop: issue the blocks due at loop step s; else every block the loop did not issue
  auto pload_at = [&](int s, bool in_loop) {
    const int nsl = K / BK;
    const int pf0 = nsl - 2 * NPF > 0 ? ((nsl - 2 * NPF) & ~1) : 0;
    static_for<FA>([&](auto a_c) {
      constexpr int a = decltype(a_c)::value;
      const bool issued = a < NPF && pf0 + 2 * a < nsl;   // the loop visits even s < nsl
      if (in_loop ? (issued && s == pf0 + 2 * a) : !issued) pload(a_c);
    });
  };

  // Slab s lives in LDS buffer s & 1. Invariant at the top of an even step s: slab s is
  // in buffer 0 (visible to all), R0 holds slab s + 1.
  const int nslab = K / BK;
  gload(0, R0);
  load_scale(bmax_in, sb_exp, sB);
  sstore(0, R0);
  if (nslab > 1) gload(1, R0);
  __syncthreads();
  for (int s = 0; s < nslab; s += 2) {
    if (s + 2 < nslab) gload(s + 2, R1);
    if constexpr (NPF > 0) {
      if (phase) pload_at(s, true);
    }
    mma(0);
    if (s + 1 >= nslab) break;
    sstore(1, R0);
    __syncthreads();
    if (s + 3 < nslab) gload(s + 3, R0);
    mma(1);
    if (s + 2 < nslab) {
      sstore(0, R1);
      __syncthreads();
    }
  }

  if constexpr (NPF > 0) {
    if (phase) pload_at(nslab, false);
  }

  // epilogue: for v_mfma_f32_16x16x32_f16 lane l, register v holds D[4 (l/16) + v][l % 16]
  // (32-bit byte offsets from the base pointers, as for the loads)
  const float unscale = pow2f(-(y_scale_exp + sb_exp));
  char* Cc = reinterpret_cast<char*>(C);
  float vmax = 0.0f;
#pragma unroll
  for (int a = 0; a < FA; ++a) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int m = m0 + wm + a * 16 + 4 * lq + v;
      const float rs = rowscale ? alpha * unscale * rowscale[m] : alpha * unscale;
      const unsigned rowoff = 4u * (unsigned)m * (unsigned)Ncols;
#pragma unroll
      for (int b = 0; b < FB; ++b) {
        if (NARROW && b >= fbw) continue;
        const unsigned n = (unsigned)band.map(n0 + wn + b * 16 + lr);
        float val = acc[a][b][v] * rs;
        if (phase) {
          const float other = __shfl_xor(val, 1, 64);
          float phx, phy;
          if constexpr (NPF > 0) {
            const float own_ph = php[a][v][b];
            const float other_ph = __shfl_xor(own_ph, 1, 64);
            phx = (n & 1u) ? other_ph : own_ph;
            phy = (n & 1u) ? own_ph : other_ph;
          } else {
            const float2 ph = *reinterpret_cast<const float2*>(Pc + rowoff + 4u * (n & ~1u));
            phx = ph.x;
            phy = ph.y;
          }
          val = (n & 1u) ? (other * phy + val * phx) : (val * phx - other * phy);
        }
        vmax = fmaxf(vmax, fabsf(val));
        *reinterpret_cast<float*>(Cc + rowoff + 4u * n) = val;
      }
    }
  }
  if (cmax_out) {
    for (int d = 32; d >= 1; d >>= 1) vmax = fmaxf(vmax, __shfl_xor(vmax, d, 64));
    if (lane == 0) sm.red[wave] = vmax;
    __syncthreads();
    if (tid == 0) {
      float m = sm.red[0];
      for (int w = 1; w < NT / 64; ++w) m = fmaxf(m, sm.red[w]);
      atomicMax(cmax_out, __float_as_uint(m));
    }
  }
}

// XCD-aware tile order: hardware places block id on XCD (id % 8); give each XCD a
// contiguous run of linear tiles (whole column panels).
FP_DEV int xcd_remap(int id, int nb) {
  return (nb % 8 == 0) ? (id % 8) * (nb / 8) + id / 8 : id;
}

// Blocks per CU the tile shape is built for: 128 x 256 tiles take 96 KiB of LDS
// (one block per CU, up to 512 VGPR + AGPR per lane), the smaller ones two.
template <int BM, int BN> struct TileOcc { static constexpr int value = (BM * BN >= 128 * 256) ? 1 : 2; };

template <int BM, int BN, int WM = 2, int WN = 2, int NPF = 0>
__global__ __launch_bounds__(WM * WN * 64, (WM * WN > 4 ? 1 : TileOcc<BM, BN>::value)) void hankel_gemm_f16x3(
    const h8* __restrict__ Yh, const h8* __restrict__ Yl, int y_scale_exp,
    const float* __restrict__ B, float* __restrict__ C, int M, int K, int Ncols, float alpha,
    const float* __restrict__ rowscale, const float* __restrict__ phase,
    const unsigned* __restrict__ bmax_in, unsigned* __restrict__ cmax_out,
    unsigned* __restrict__ reset_slot) {
  __shared__ Smem<BM, BN, WM, WN> sm;
  if (reset_slot && blockIdx.x == 0 && threadIdx.x == 0) *reset_slot = 0u;
  const int tilesM = M / BM, tilesN = Ncols / BN;
  const int lin = xcd_remap(blockIdx.x, tilesM * tilesN);
  gemm_tile<BM, BN, WM, WN, NPF>(sm, lin % tilesM, lin / tilesM, Yh, Yl, y_scale_exp, B, C, M, K,
                                 Ncols, alpha, rowscale, phase, bmax_in, cmax_out);
}

// Single-instance launch restricted to a column band: tiles of 256 rows x TBK 16-column
// blocks (8 waves, one block per CU) over the virtual columns of `band`, `slots` tiles per
// row panel. Afterwards every block writes zeros to its share of the real columns
// [zero_lo, zero_hi) (the masked frequencies; empty range: no writes).
template <int TBK, int NPF>
__global__ __launch_bounds__(512, 1) void hankel_gemm_f16x3_band(
    const h8* __restrict__ Yh, const h8* __restrict__ Yl, int y_scale_exp,
    const float* __restrict__ B, float* __restrict__ C, int M, int K, int Ncols, float alpha,
    const float* __restrict__ rowscale, const float* __restrict__ phase,
    const unsigned* __restrict__ bmax_in, unsigned* __restrict__ cmax_out,
    unsigned* __restrict__ reset_slot, int slots, ColBand band, int zero_lo, int zero_hi) {
  __shared__ Smem<256, 128, 4, 2> sm;
  if (reset_slot && blockIdx.x == 0 && threadIdx.x == 0) *reset_slot = 0u;
  const int tilesM = M / 256;
  const int nb = tilesM * slots;
  const int lin = xcd_remap(blockIdx.x, nb);
  const int tm = lin % tilesM, sl = lin / tilesM;
  gemm_tile<256, 128, 4, 2, NPF, TBK>(sm, tm, 0, Yh, Yl, y_scale_exp, B, C, M, K, Ncols, alpha,
                                      rowscale, phase, bmax_in, cmax_out, band,
                                      sl * TBK * 16);
  if (zero_hi > zero_lo) {
    const int w4 = (zero_hi - zero_lo) / 4;
    const long long total = (long long)M * w4;
    const long long chunk = (total + gridDim.x - 1) / gridDim.x;
    const long long end = min(total, (long long)(blockIdx.x + 1) * chunk);
    for (long long i = (long long)blockIdx.x * chunk + threadIdx.x; i < end; i += 512) {
      const int r = (int)(i / w4), c4 = (int)(i % w4);
      *reinterpret_cast<float4*>(C + (size_t)r * Ncols + zero_lo + 4 * c4) =
          make_float4(0.f, 0.f, 0.f, 0.f);
    }
  }
}


// Band GEMM without a phase (G2 / G4 of the z-step), K split across two wave groups.
// Tile: 256 rows x TBK 16-column blocks, 8 waves. Waves 0-3 (group 0) multiply the even K
// slabs, waves 4-7 (group 1) the odd ones; wave w owns rows m0 + 64 (w % 4) of its group's
// partial sum (all TBK blocks: 4 x TBK accumulators). Each wave loads its own A fragments
// (Y is stored in fragment order, 1 KB contiguous per fragment) straight from global memory
// into registers two slabs ahead, so A never passes through LDS and no A fragment is loaded
// twice per tile; each group splits its slabs' B into LDS (double-buffered per group) and
// four waves read it. The two partial sums are added through LDS at the end: group 0 stores
// blocks [0, ceil(TBK / 2)), group 1 the rest. The accumulation order differs from
// hankel_gemm_f16x3_band (two interleaved partial sums over K), so the results agree with
// it to FP32 rounding, not bitwise. Needs K % 128 == 0.
template <int TBK> struct SmemKSplit {
  static constexpr int RED = 4 * TBK * 4 * 4 * 64;   // partial sums [wr][b][a][v][lane]
  union {
    struct {
      h8 Bh[2][2][128 * 4];   // [group][buffer][fragment]
      h8 Bl[2][2][128 * 4];
    } b;
    float red[RED];
  } u;
  float wmax[8];
};

template <int TBK>
__global__ __launch_bounds__(512, 1) void hankel_gemm_f16x3_ksplit(
    const h8* __restrict__ Yh, const h8* __restrict__ Yl, int y_scale_exp,
    const float* __restrict__ B, float* __restrict__ C, int M, int K, int Ncols, float alpha,
    const float* __restrict__ rowscale, const unsigned* __restrict__ bmax_in,
    unsigned* __restrict__ cmax_out, unsigned* __restrict__ reset_slot, int slots, ColBand band,
    int zero_lo, int zero_hi) {
  __shared__ SmemKSplit<TBK> sm;
  constexpr int FA = 4;                        // 16-row blocks per wave
  constexpr int NB0 = (TBK + 1) / 2;           // blocks stored by group 0
  if (reset_slot && blockIdx.x == 0 && threadIdx.x == 0) *reset_slot = 0u;
  const int tilesM = M / 256;
  const int lin = xcd_remap(blockIdx.x, tilesM * slots);
  const int tm = lin % tilesM, sl = lin / tilesM;
  const int m0 = tm * 256, n0 = sl * TBK * 16;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int g = wave >> 2;                     // K group: slabs g, g + 2, g + 4, ...
  const int wr = wave & 3;                     // wave row
  const int tg = tid & 255;                    // thread within the group
  const int lq = lane >> 4, lr = lane & 15;

  // B staging: 2 threads per column, each 2 k groups of 8 rows; columns past the tile's
  // TBK blocks load its last column (staged, never multiplied)
  const int bc = tg & 127;
  const int bkp = tg >> 7;
  const char* Bc = reinterpret_cast<const char*>(B);
  const unsigned bcol4 = 4u * (unsigned)band.map(n0 + (bc < TBK * 16 ? bc : TBK * 16 - 1));
  const unsigned row4 = 4u * (unsigned)Ncols;
  // A fragment (mb, ks) of this wave: h8 entry (ks (M / 16) + mb) 64 + lane
  const char* Yhc = reinterpret_cast<const char*>(Yh);
  const char* Ylc = reinterpret_cast<const char*>(Yl);
  const unsigned a_slab = 16u * 64u * (unsigned)(M / 16);
  const unsigned a_base = 16u * (64u * (unsigned)(m0 / 16 + FA * wr) + (unsigned)lane);

  struct AFrag {
    h8 h[FA], l[FA];
  };
  auto gload_a = [&](int ks, AFrag& A) {
    const unsigned off = (unsigned)ks * a_slab + a_base;
#pragma unroll
    for (int a = 0; a < FA; ++a) {
      A.h[a] = *reinterpret_cast<const h8*>(Yhc + off + 1024u * a);
      A.l[a] = *reinterpret_cast<const h8*>(Ylc + off + 1024u * a);
    }
  };
  struct BStage {
    float v[16];
  };
  auto gload_b = [&](int ks, BStage& R) {
    const unsigned k0 = (unsigned)(ks * BK);
#pragma unroll
    for (int gg = 0; gg < 2; ++gg) {
      const unsigned r0 = (k0 + 8u * (unsigned)(2 * bkp + gg)) * row4 + bcol4;
#pragma unroll
      for (int j = 0; j < 8; ++j)
        R.v[gg * 8 + j] = *reinterpret_cast<const float*>(Bc + (r0 + (unsigned)j * row4));
    }
  };
  int sb_exp = 0;
  float sB = 1.0f;
  auto sstore_b = [&](int buf, const BStage& R) {
#pragma unroll
    for (int gg = 0; gg < 2; ++gg) {
      const int kg = 2 * bkp + gg;
      h8 hi, lo;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const float x = R.v[gg * 8 + j] * sB;
        const _Float16 h = (_Float16)x;
        hi[j] = h;
        lo[j] = (_Float16)(x - (float)h);
      }
      const int idx = ((bc >> 4) * 4 + kg) * 16 + (bc & 15);
      sm.u.b.Bh[g][buf][idx] = hi;
      sm.u.b.Bl[g][buf][idx] = lo;
    }
  };

  f4 acc[FA][TBK];
#pragma unroll
  for (int a = 0; a < FA; ++a)
#pragma unroll
    for (int b = 0; b < TBK; ++b) acc[a][b] = f4{0.f, 0.f, 0.f, 0.f};

  // Per column block the FA products Yl Bh, then the FA Yh Bl, then the FA Yh Bh: each
  // accumulator gets its three updates in the same order as in gemm_tile, with FA - 1
  // independent MFMAs between two of them. (Measured against one accumulator at a time,
  // the raised wave priority and the scheduling groups of gemm_tile: 46 vs 50 us at
  // 1024 x 6240 band columns; those help only when A also comes through LDS.)
  auto mma = [&](int buf, const AFrag& A) {
    const h8* bh_s = sm.u.b.Bh[g][buf];
    const h8* bl_s = sm.u.b.Bl[g][buf];
#pragma unroll
    for (int b = 0; b < TBK; ++b) {
      const int idx = (b * 4 + lq) * 16 + lr;
      const h8 bh = bh_s[idx];
      const h8 bl = bl_s[idx];
#pragma unroll
      for (int a = 0; a < FA; ++a)
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A.l[a], bh, acc[a][b], 0, 0, 0);
#pragma unroll
      for (int a = 0; a < FA; ++a)
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A.h[a], bl, acc[a][b], 0, 0, 0);
#pragma unroll
      for (int a = 0; a < FA; ++a)
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A.h[a], bh, acc[a][b], 0, 0, 0);
    }
  };

  // Local step t multiplies the group's slab g + 2 t from LDS buffer t & 1 with A set t & 1.
  // Invariant at the top of an even step: buffer 0 holds step t's B, A0 / A1 steps t and
  // t + 1, R0 step t + 1's B. T = K / 64 local steps, even.
  const int T = K / BK / 2;
  AFrag A0, A1;
  BStage R0;
  gload_b(g, R0);
  gload_a(g, A0);
  load_scale(bmax_in, sb_exp, sB);
  sstore_b(0, R0);
  gload_b(g + 2, R0);
  gload_a(g + 2, A1);
  __syncthreads();
  for (int t = 0; t < T - 2; t += 2) {
    mma(0, A0);
    gload_a(g + 2 * (t + 2), A0);
    sstore_b(1, R0);
    gload_b(g + 2 * (t + 2), R0);
    __syncthreads();
    mma(1, A1);
    gload_a(g + 2 * (t + 3), A1);
    sstore_b(0, R0);
    gload_b(g + 2 * (t + 3), R0);
    __syncthreads();
  }
  mma(0, A0);
  sstore_b(1, R0);
  __syncthreads();
  mma(1, A1);

  // add the two partial sums: each group sends the blocks the other one stores
  float* red = sm.u.red;
  auto ridx = [&](int b, int a, int v) { return (((wr * TBK + b) * FA + a) * 4 + v) * 64 + lane; };
  __syncthreads();
#pragma unroll
  for (int b = 0; b < TBK; ++b) {
    if ((g == 0) == (b < NB0)) continue;
#pragma unroll
    for (int a = 0; a < FA; ++a)
#pragma unroll
      for (int v = 0; v < 4; ++v) red[ridx(b, a, v)] = acc[a][b][v];
  }
  __syncthreads();

  // epilogue: for v_mfma_f32_16x16x32_f16 lane l, register v holds D[4 (l/16) + v][l % 16]
  const float unscale = pow2f(-(y_scale_exp + sb_exp));
  char* Cc = reinterpret_cast<char*>(C);
  float vmax = 0.0f;
#pragma unroll
  for (int a = 0; a < FA; ++a) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int m = m0 + 64 * wr + a * 16 + 4 * lq + v;
      const float rs = rowscale ? alpha * unscale * rowscale[m] : alpha * unscale;
      const unsigned rowoff = 4u * (unsigned)m * (unsigned)Ncols;
#pragma unroll
      for (int b = 0; b < TBK; ++b) {
        if ((g == 0) != (b < NB0)) continue;
        const unsigned n = (unsigned)band.map(n0 + b * 16 + lr);
        const float val = (acc[a][b][v] + red[ridx(b, a, v)]) * rs;
        vmax = fmaxf(vmax, fabsf(val));
        *reinterpret_cast<float*>(Cc + rowoff + 4u * n) = val;
      }
    }
  }
  if (cmax_out) {
    for (int d = 32; d >= 1; d >>= 1) vmax = fmaxf(vmax, __shfl_xor(vmax, d, 64));
    if (lane == 0) sm.wmax[wave] = vmax;
    __syncthreads();
    if (tid == 0) {
      float mm = sm.wmax[0];
      for (int w = 1; w < 8; ++w) mm = fmaxf(mm, sm.wmax[w]);
      atomicMax(cmax_out, __float_as_uint(mm));
    }
  }
  if (zero_hi > zero_lo) {
    const int w4 = (zero_hi - zero_lo) / 4;
    const long long total = (long long)M * w4;
    const long long chunk = (total + gridDim.x - 1) / gridDim.x;
    const long long end = min(total, (long long)(blockIdx.x + 1) * chunk);
    for (long long i = (long long)blockIdx.x * chunk + threadIdx.x; i < end; i += 512) {
      const int r = (int)(i / w4), c4 = (int)(i % w4);
      *reinterpret_cast<float4*>(C + (size_t)r * Ncols + zero_lo + 4 * c4) =
          make_float4(0.f, 0.f, 0.f, 0.f);
    }
  }
}

}  // namespace fpgemm16
