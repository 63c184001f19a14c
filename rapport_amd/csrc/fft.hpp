// LDS-resident Stockham FFT along t for one row held by one workgroup.
//
// Layout: a workgroup of TPB threads owns one row of N complex samples; thread
// `tid` keeps E = N/TPB samples in registers in the *stride layout*
// v[s] = x[tid + s*TPB]. Each radix-R pass does E/R butterflies per thread on
// its own registers (butterfly j = tid + q*TPB reads x[j + r*N/R], which is
// register q + r*E/R), writes the Stockham output positions to LDS, and the
// next pass reads the stride layout back. The last pass writes x[j + r*N/R]
// again, so input and output share the stride layout and a row can go
// forward -> pointwise multiply -> inverse without leaving registers.
//
// Conventions are numpy's: forward X[k] = sum_n x[n] exp(-2 pi i k n / N),
// inverse has +i and is NOT scaled here (callers fold the 1/N).
//
// Twiddles come from a per-N table tw[m] = exp(-2 pi i m / N), m < N. The exact float
// variant (EXACT, E = 4, 8 or 16 samples per thread) reads compact per-pass tables stored
// after tw[2 N] instead (exact_table_offset).
#pragma once
#include "common.hpp"

namespace fpfft {

// LDS slot of element i of the exchange buffer: an XOR swizzle of the low five bits by
// bits 3..7, and of bit 0 by bit 7 (only the low five bits change, each from higher bits,
// so it is a bijection on [0, N)). It is linear over XOR: lds_pad(a ^ b) = lds_pad(a) ^
// lds_pad(b). With the LDS serving 16 lanes of 8-byte accesses (8 of 16-byte ones) per
// cycle from 32 banks, every Stockham write and stride-layout read of the complex64
// transforms at 8 or 16 samples per thread, and of the complex128 ones at 8, is free of
// bank conflicts (checked on the host for every row shape; at 16 samples per thread the
// writes of 16 lanes span index bits 4..7, which bit 7 -> bit 0 makes distinct).
__host__ __device__ constexpr inline int lds_pad(int i) {
  return i ^ ((i >> 3) & 31) ^ ((i >> 7) & 1);
}
// The XOR mask lds_pad applies to i (a constant when i is).
__host__ __device__ constexpr inline int lds_mask(int i) { return lds_pad(i) ^ i; }
template <int N> struct LdsSize { static constexpr int value = N; };

constexpr int ilog2(int v) { return v <= 1 ? 0 : 1 + ilog2(v >> 1); }
constexpr int bitrev(int r, int nbits) {
  int o = 0;
  for (int b = 0; b < nbits; ++b) o |= ((r >> b) & 1) << (nbits - 1 - b);
  return o;
}

// cos / sin of 2*pi*k/16
template <typename T> FP_DEV T c16(int k) {
  constexpr double tab[8] = {1.0, 0.92387953251128675613, 0.70710678118654752440,
                             0.38268343236508977173, 0.0, -0.38268343236508977173,
                             -0.70710678118654752440, -0.92387953251128675613};
  return T(tab[k & 7]);
}
template <typename T> FP_DEV T s16(int k) {
  constexpr double tab[8] = {0.0, 0.38268343236508977173, 0.70710678118654752440,
                             0.92387953251128675613, 1.0, 0.92387953251128675613,
                             0.70710678118654752440, 0.38268343236508977173};
  return T(tab[k & 7]);
}

// The float roundings of cos / sin(2 pi k / 16) all lie inside the unit circle
// (|(c, s)| - 1 = -1.7e-8 for k = 2, 6 and -2.9e-8 for odd k), so every rotation by them
// shrinks its operand: a coherent loss of about 6e-8 in amplitude per float transform.
// Their residuals (cos - float(cos), likewise sin) let the exact variant below form each
// product to within the rounding of its result.
template <typename T> FP_DEV T c16lo(int k) {
  constexpr double tab[8] = {0.0, 0.92387953251128675613, 0.70710678118654752440,
                             0.38268343236508977173, 0.0, -0.38268343236508977173,
                             -0.70710678118654752440, -0.92387953251128675613};
  return T(tab[k & 7] - (double)(T)tab[k & 7]);
}
template <typename T> FP_DEV T s16lo(int k) {
  constexpr double tab[8] = {0.0, 0.38268343236508977173, 0.70710678118654752440,
                             0.92387953251128675613, 1.0, 0.92387953251128675613,
                             0.70710678118654752440, 0.38268343236508977173};
  return T(tab[k & 7] - (double)(T)tab[k & 7]);
}

// x * exp(-+2 pi i idx/16), idx in [0, 8). EXACT (float only): products with the
// constant split in a rounded part and its residual, fused, so the rotation has no
// magnitude bias.
template <typename T, bool INV, bool EXACT = false> FP_DEV cpx<T> rot16(cpx<T> x, int idx) {
  if (idx == 0) return x;
  if (idx == 4) return mul_mi<T, INV>(x);
  const T c = c16<T>(idx), s = INV ? s16<T>(idx) : -s16<T>(idx);
  if constexpr (EXACT && sizeof(T) == 4) {
    const T cl = c16lo<T>(idx), sl = INV ? s16lo<T>(idx) : -s16lo<T>(idx);
    const T re_lo = __builtin_fmaf(x.x, cl, -x.y * sl);
    const T im_lo = __builtin_fmaf(x.x, sl, x.y * cl);
    return mk<T>(__builtin_fmaf(x.x, c, __builtin_fmaf(-x.y, s, re_lo)),
                 __builtin_fmaf(x.x, s, __builtin_fmaf(x.y, c, im_lo)));
  }
  return mk<T>(x.x * c - x.y * s, x.x * s + x.y * c);
}

// In-register DFT of size R (R in {2,4,8,16}), natural order in and out.
template <typename T, int R, bool INV, bool EXACT = false> FP_DEV void dft(cpx<T> (&a)[R]) {
  constexpr int LB = ilog2(R);
#pragma unroll
  for (int len = R; len >= 2; len >>= 1) {
    const int half = len >> 1;
#pragma unroll
    for (int st = 0; st < R; st += len) {
#pragma unroll
      for (int m = 0; m < half; ++m) {
        cpx<T> x = a[st + m], y = a[st + m + half];
        a[st + m] = x + y;
        a[st + m + half] = rot16<T, INV, EXACT>(x - y, m * (R / len) * (16 / R));
      }
    }
  }
  cpx<T> b[R];
#pragma unroll
  for (int r = 0; r < R; ++r) b[r] = a[bitrev(r, LB)];
#pragma unroll
  for (int r = 0; r < R; ++r) a[r] = b[r];
}

// Compact twiddle tables of the exact float transform, stored after the 2 N entries of
// the twiddle table: for a row held E samples per thread the passes are radix
// R = min(E, N / Ns) at Ns = 1, R, R R', ...; each twiddled pass (Ns > 1) stores, for
// r = 1 .. R-1 and k < Ns (entry (r - 1) Ns + k), the float4 (w^r rounded, residual),
// w = exp(-2 pi i k N / (Ns R) / N).
// The table for E = 8 comes first, then the one for E = 4, then the one for E = 16.
constexpr int exact_pass_entries(int N, int E, int Ns) {
  return Ns >= N ? 0
                 : (Ns > 1 ? Ns * ((N / Ns >= E ? E : N / Ns) - 1) : 0) +
                       exact_pass_entries(N, E, Ns * (N / Ns >= E ? E : N / Ns));
}
template <int N, int E, int NsTarget>
constexpr int exact_table_offset() {
  int off = E == 8 ? 0 : exact_pass_entries(N, 8, 1);
  if (E == 16) off += exact_pass_entries(N, 4, 1);
  for (int ns = 1; ns < NsTarget;) {
    const int r = N / ns >= E ? E : N / ns;
    if (ns > 1) off += ns * (r - 1);
    ns *= r;
  }
  return off;
}

// FP64, one butterfly per thread and pass (E = R): the table twiddle w^1 of every pass is
// loaded when the transform starts (pre[d] for the pass of depth d), not at the top of its
// pass, so the loads overlap each other and the first passes instead of stalling each pass
// (the FP64 row kernels run at two waves per SIMD).
template <typename T, int N, int TPB>
struct TwPre {
  static constexpr int kMaxPasses = 16;   // log2 N passes at most (radix 2, N <= 65536)
  cpx<T> w[kMaxPasses];
};

template <typename T, int N, int TPB, int Ns, int D>
FP_DEV void prefetch_twiddles(TwPre<T, N, TPB>& pre, const cpx<T>* __restrict__ tw, int tid) {
  constexpr int E = N / TPB;
  constexpr int REM = N / Ns;
  constexpr int R = REM >= E ? E : REM;
  static_assert(D < TwPre<T, N, TPB>::kMaxPasses, "more transform passes than prefetch slots");
  if constexpr (Ns > 1 && E / R == 1) pre.w[D] = tw[(tid & (Ns - 1)) * (N / (Ns * R))];
  if constexpr (Ns * R < N) prefetch_twiddles<T, N, TPB, Ns * R, D + 1>(pre, tw, tid);
}

template <typename T, int N, int TPB, bool INV, int Ns, bool EXACT = false, int D = 0>
FP_DEV void passes(cpx<T> (&v)[N / TPB], cpx<T>* lds, const cpx<T>* __restrict__ tw, int tid,
                   const TwPre<T, N, TPB>* pre = nullptr) {
  constexpr int E = N / TPB;
  constexpr int REM = N / Ns;
  constexpr int R = REM >= E ? E : REM;
  constexpr bool LAST = (Ns * R == N);
  constexpr int Q = E / R;
  constexpr int STRIDE = N / (Ns * R);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int j = tid + q * TPB;
    const int k = j & (Ns - 1);
    cpx<T> a[R];
#pragma unroll
    for (int r = 0; r < R; ++r) a[r] = v[q + r * Q];
    if (Ns > 1) {
      if (EXACT && sizeof(T) == 4) {
        // every power from the pass's compact table (exact_table_offset): one float4
        // (rounded w^r, residual) per power, stored power-major ([r - 1][k]) so the lanes
        // of a wave, whose k are consecutive (or repeat), read one contiguous run per
        // power; products fused with the residual: no rounding of the twiddles is raised
        // to a power, and the transform has no coherent gain
        const float4* pt = reinterpret_cast<const float4*>(tw + 2 * N) +
                           exact_table_offset<N, E, Ns>() + k;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          const float4 e = pt[(r - 1) * Ns];
          const T hy = INV ? -e.y : e.y, ly = INV ? -e.w : e.w;
          const cpx<T> x = a[r];
          const T re_lo = __builtin_fmaf(x.x, e.z, -x.y * ly);
          const T im_lo = __builtin_fmaf(x.x, ly, x.y * e.z);
          a[r] = mk<T>(__builtin_fmaf(x.x, e.x, __builtin_fmaf(-x.y, hy, re_lo)),
                       __builtin_fmaf(x.x, hy, __builtin_fmaf(x.y, e.x, im_lo)));
        }
      } else if (sizeof(T) == 8) {
        // FP64: powers of one table twiddle (rounding growth ~R ulp is below FP64 needs)
        cpx<T> w1 = (Q == 1 && pre) ? pre->w[D] : tw[k * STRIDE];
        if (INV) w1 = conjx(w1);
        cpx<T> w = w1;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          a[r] = a[r] * w;
          if (r + 1 < R) w = w * w1;
        }
      } else {
        // FP32: one table load, w^1, per butterfly; the other powers by products
        // (w^2 = w w, w^4 = w^2 w^2, w^8 = w^4 w^4, then one product each): at most
        // 7 x 2^-24 from exp(-2 pi i m / N) for R = 8 and 15 x 2^-24 for R = 16
        // (N = 4096). The table loads are scattered (lane stride STRIDE) and cost
        // more than the products: loading w^4 (and w^8) as well is 2.7 % slower per
        // mixed z-step at 1024 x 4096, loading all R - 1 powers 25 % slower.
        cpx<T> w[R];
        w[0] = mk<T>(T(1), T(0));
        w[1] = tw[k * STRIDE];
        if (R >= 4) {
          w[2] = w[1] * w[1];
          w[3] = w[2] * w[1];
        }
        if (R >= 8) {
          w[4] = w[2] * w[2];
          w[5] = w[4] * w[1];
          w[6] = w[4] * w[2];
          w[7] = w[4] * w[3];
        }
        if (R >= 16) {
          w[8] = w[4] * w[4];
#pragma unroll
          for (int r = 9; r < 16; ++r) w[r] = w[8] * w[r - 8];
        }
#pragma unroll
        for (int r = 1; r < R; ++r) a[r] = a[r] * (INV ? conjx(w[r]) : w[r]);
      }
    }
    dft<T, R, INV, EXACT>(a);
    if (LAST) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[q + r * Q] = a[r];
    } else {
      // idxD and r Ns occupy disjoint bits (so their sum is their XOR) and lds_pad is
      // linear over XOR, so lds_pad(idxD + r Ns) = lds_pad(idxD) ^ lds_pad(r Ns): one
      // swizzled base per butterfly and one XOR with a constant per output
      // (byte offsets: the element size is a power of two, so the XOR commutes with it)
      const int idxD = (j / Ns) * Ns * R + k;
      constexpr unsigned ES = sizeof(cpx<T>);
      const unsigned bD = (unsigned)lds_pad(idxD) * ES;
      char* lb = reinterpret_cast<char*>(lds);
#pragma unroll
      for (int r = 0; r < R; ++r)
        *reinterpret_cast<cpx<T>*>(lb + (bD ^ ((unsigned)lds_pad(r * Ns) * ES))) = a[r];
    }
  }
  if constexpr (!LAST) {
    __syncthreads();
    if constexpr (TPB % 256 == 0) {
      // lds_pad(tid + s TPB) = (lds_pad(tid) ^ lds_mask(s TPB)) + s TPB, since lds_pad(tid)
      // and the mask stay below TPB and s TPB has no bits below it; lds_pad reads no bit
      // above 7, so the mask is 0: one swizzled base per thread, the rest immediate offsets
      const int b = lds_pad(tid);
#pragma unroll
      for (int s = 0; s < E; ++s) v[s] = lds[(b ^ lds_mask(s * TPB)) + s * TPB];
    } else {
#pragma unroll
      for (int s = 0; s < E; ++s) v[s] = lds[lds_pad(tid + s * TPB)];
    }
    __syncthreads();
    passes<T, N, TPB, INV, Ns * R, EXACT, D + 1>(v, lds, tw, tid, pre);
  }
}

// Full-row FFT in the stride layout. `lds` must hold LdsSize<N>::value entries
// and must not be in use by other threads of the workgroup when called.
template <typename T, int N, int TPB, bool INV, bool EXACT = false>
FP_DEV void row_fft(cpx<T> (&v)[N / TPB], cpx<T>* lds, const cpx<T>* __restrict__ tw) {
  static_assert(N % TPB == 0, "N must be a multiple of TPB");
  static_assert(!EXACT || N / TPB == 4 || N / TPB == 8 || N / TPB == 16,
                "exact transform: 4, 8 or 16 samples per thread");
  // Opaque copy of the thread index: keeps the compiler from hoisting this
  // transform's LDS / twiddle addresses out of enclosing loops and holding
  // them in registers across the whole kernel.
  int tid = (int)threadIdx.x & (TPB - 1);
  asm volatile("" : "+v"(tid));
  if constexpr (sizeof(T) == 8 && !EXACT) {
    TwPre<T, N, TPB> pre;
    prefetch_twiddles<T, N, TPB, 1, 0>(pre, tw, tid);
    passes<T, N, TPB, INV, 1, EXACT>(v, lds, tw, tid, &pre);
  } else {
    passes<T, N, TPB, INV, 1, EXACT>(v, lds, tw, tid);
  }
}

}  // namespace fpfft
