// Nonlinear operator of the split step in FP32 (mixed-precision path), one
// workgroup per radial row: the same arithmetic as nl.hpp (split() of
// D3/sim/operators.py:89-146 and the nonlinear part of Integrator.step,
// D3/sim/integrator.py:141-147) with complex64 rows, FP32 transforms and an
// FP32 PCHIP. The carrier densities stay FP64 in memory and are rounded when
// read. The field-loss PCHIP is evaluated in the local coordinate
// s = (I - x_i)/h_i with coefficients pre-multiplied by photo_pref, so every
// intermediate stays inside FP32 range (|W photo_pref| < 6e30 over the table).
#pragma once
#include "common.hpp"
#include "fft.hpp"
#include "raman.hpp"
#include "../../include/rapport_amd.h"

namespace fpnl32 {

using fpfft::LdsSize;
using fpfft::row_fft;
typedef cpx<float> c32;

// (10^k, 10^-k) as floats, k = 0 .. 17: the decade boundaries of the knot grid and their
// reciprocals.
__device__ const float2 kDec[18] = {
    {1e0f, 1e0f},     {1e1f, 1e-1f},   {1e2f, 1e-2f},   {1e3f, 1e-3f},   {1e4f, 1e-4f},
    {1e5f, 1e-5f},    {1e6f, 1e-6f},   {1e7f, 1e-7f},   {1e8f, 1e-8f},   {1e9f, 1e-9f},
    {1e10f, 1e-10f},  {1e11f, 1e-11f}, {1e12f, 1e-12f}, {1e13f, 1e-13f}, {1e14f, 1e-14f},
    {1e15f, 1e-15f},  {1e16f, 1e-16f}, {1e17f, 1e-17f}};

FP_DEV float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// PCHIP interval of the (clamped) intensity x on the knot grid 10^k (1 + 0.01 j):
// returns the coefficient index i, the local coordinate s in [0, 1] and 1 / 10^k.
// Branch-free: the binary exponent gives the decade k0 = floor(e2 log10 2), which is
// floor(log10 x) or one below it; both candidate table entries are loaded at once and
// the comparison with 10^(k0 + 1) picks one. (With a branch per sample the compiler
// waited for each table load in turn, eight samples one after another.)
FP_DEV int pchip_piece(float x, int n, float& s, float& inv_pk) {
  const int e2 = (int)((__float_as_uint(x) >> 23) & 0xff) - 127;
  int k = e2 <= 0 ? 0 : (e2 * 1233) >> 12;   // floor(e2 log10 2), possibly one low
  if (k > 16) k = 16;
  const float2 d0 = kDec[k], d1 = kDec[k + 1];
  const bool up = k < 16 && x >= d1.x;
  k += up ? 1 : 0;
  inv_pk = up ? d1.y : d0.y;
  const float v = (x * inv_pk - 1.0f) * 100.0f;
  int j = (int)v;
  j = j < 0 ? 0 : (j > 899 ? 899 : j);
  int i = k * 900 + j;
  float sl = v - (float)j;
  const bool last = i > n - 2;
  i = last ? n - 2 : i;
  sl = last ? 1.0f : sl;
  s = fminf(fmaxf(sl, 0.0f), 1.0f);
  return i;
}

// Coefficients (c0, c1, c2, c3) of PCHIP interval i: one 16-byte load from the
// interleaved (n_knots - 1) x 4 table.
FP_DEV float4 pchip_coef(const fp_nl_tables_f32& t, int i) {
  return reinterpret_cast<const float4*>(t.pcoef_s)[i];
}

// |W(I)| * photo_pref * depl / (|u|^2 + 1e-30), the photo-ionisation loss rate.
// The knot grid is 10^k (1 + 0.01 j), so the interval and the local coordinate
// follow from the decade alone: v = (I / 10^k - 1) * 100, j = floor(v),
// s = v - j. Only the four coefficients of interval k*900 + j are loaded.
FP_DEV float photo_rate(float a2, float rho, const fp_nl_params& p, const fp_nl_tables_f32& t) {
  float x = a2 * (float)p.invE2;
  x = fminf(fmaxf(x, (float)p.Imin), (float)p.Imax);
  const int n = p.n_knots;
  float sl, ipk;
  const int i = pchip_piece(x, n, sl, ipk);
  const float4 c = pchip_coef(t, i);
  const float w = fabsf(((c.x * sl + c.y) * sl + c.z) * sl + c.w);
  float depl = 1.0f - rho * rcp((float)p.rho_max);
  depl = fminf(fmaxf(depl, 0.0f), 1.0f);
  return nan0(w) * depl * rcp(a2 + 1e-30f);
}

// rho_row: the row's electron density as float (LDS in the fused step);
// rhos_row: the STE density row in global memory (only read when RS).
template <int NT, int TPB, bool RS>
FP_DEV void split_row(const c32 (&us)[NT / TPB], const float* rho_row,
                      const double* __restrict__ rhos_row, c32 (&k)[NT / TPB], float* al,
                      const fp_nl_params& p, const fp_nl_tables_f32& t, c32* lds) {
  constexpr int E = NT / TPB;
  int tid = threadIdx.x & (TPB - 1);
  asm volatile("" : "+v"(tid));
  const c32* tw = reinterpret_cast<const c32*>(t.tw);
  const c32* Rf = reinterpret_cast<const c32*>(t.Rf);
  const float invN = 1.0f / (float)NT;
  c32 buf[E];
  float phs[E];                        // photo-ionisation rate, reused for alpha below
  if (p.en_pi_loss) {
#pragma unroll
    for (int s = 0; s < E; ++s) {
      const float a2 = us[s].x * us[s].x + us[s].y * us[s].y;
      phs[s] = photo_rate(a2, rho_row[tid + s * TPB], p, t);
      buf[s] = phs[s] * us[s];
    }
    row_fft<float, NT, TPB, false>(buf, lds, tw);
#pragma unroll
    for (int s = 0; s < E; ++s) k[s] = t.mB[tid + s * TPB] * buf[s];
  } else {
#pragma unroll
    for (int s = 0; s < E; ++s) {
      k[s] = mk<float>(0.f, 0.f);
      phs[s] = 0.0f;
    }
  }
  if (p.en_kerr_inst || p.en_kerr_raman) {
    float KR[E];
    if (p.en_kerr_raman)
      fpraman::raman_row<float, NT, TPB>(us, KR, lds, reinterpret_cast<const c32*>(t.rw));
    const float fR = (float)p.f_R, kp = (float)p.kerr_pref;
#pragma unroll
    for (int s = 0; s < E; ++s) {
      const float a2 = us[s].x * us[s].x + us[s].y * us[s].y;
      float kerr = p.en_kerr_inst ? (1.0f - fR) * a2 : 0.0f;
      if (p.en_kerr_raman) kerr = kerr + fR * KR[s];
      const float kk = kp * kerr;
      buf[s] = mk<float>(-kk * us[s].y, kk * us[s].x);
    }
    row_fft<float, NT, TPB, false>(buf, lds, tw);
#pragma unroll
    for (int s = 0; s < E; ++s) k[s] = k[s] + t.mA[tid + s * TPB] * buf[s];
  }
  const bool any_extra = (p.pc_re != 0.0 || p.pc_im != 0.0) || (RS && p.ste_pref != 0.0);
  if (any_extra) {
    const c32 g = mk<float>((float)(-p.plasma_pref * p.pc_re), (float)(-p.plasma_pref * p.pc_im));
#pragma unroll
    for (int s = 0; s < E; ++s) {
      c32 e = rho_row[tid + s * TPB] * (g * us[s]);
      if (RS && p.ste_pref != 0.0) {
        const float q = (float)(p.ste_pref * rhos_row[tid + s * TPB]);
        e = e + mk<float>(-q * us[s].y, q * us[s].x);
      }
      buf[s] = e;
    }
    row_fft<float, NT, TPB, false>(buf, lds, tw);
#pragma unroll
    for (int s = 0; s < E; ++s) k[s] = k[s] + t.mC[tid + s * TPB] * buf[s];
  }
  row_fft<float, NT, TPB, true>(k, lds, tw);
#pragma unroll
  for (int s = 0; s < E; ++s) {
    const float a2 = us[s].x * us[s].x + us[s].y * us[s].y;
    float a = phs[s];
    if (p.en_plasma_absorb) a = a + (float)p.plasma_pref * rho_row[tid + s * TPB];
    a = nan0(a);
    a = a > 0.0f ? a : 0.0f;
    if (al) al[s] = a;
    c32 r = mk<float>(k[s].x * invN + a * us[s].x, k[s].y * invN + a * us[s].y);
    if (!(a2 >= 1e-30f)) r = mk<float>(0.f, 0.f);
    k[s] = mk<float>(nan0(r.x), nan0(r.y));
  }
}

template <int NT, int TPB>
FP_DEV void load_row(const float* __restrict__ src, size_t row, c32 (&v)[NT / TPB]) {
  const c32* r = reinterpret_cast<const c32*>(src) + row * NT;
#pragma unroll
  for (int s = 0; s < NT / TPB; ++s) v[s] = r[(threadIdx.x & (TPB - 1)) + s * TPB];
}
template <int NT, int TPB>
FP_DEV void store_row(float* __restrict__ dst, size_t row, const c32 (&v)[NT / TPB]) {
  c32* r = reinterpret_cast<c32*>(dst) + row * NT;
#pragma unroll
  for (int s = 0; s < NT / TPB; ++s) r[(threadIdx.x & (TPB - 1)) + s * TPB] = v[s];
}

template <int NT, int TPB, bool RS>
__global__ __launch_bounds__(TPB) void split_kernel(const float* __restrict__ u,
                                                     const double* __restrict__ rho,
                                                     const double* __restrict__ rhos,
                                                     float* __restrict__ rhs,
                                                     float* __restrict__ alpha, fp_nl_params p,
                                                     fp_nl_tables_f32 t) {
  __shared__ c32 lds[LdsSize<NT>::value];
  __shared__ float sr[NT];
  constexpr int E = NT / TPB;
  const size_t row = blockIdx.x;
  c32 us[E], k[E];
  float al[E];
  load_row<NT, TPB>(u, row, us);
#pragma unroll
  for (int s = 0; s < E; ++s) sr[(threadIdx.x & (TPB - 1)) + s * TPB] = (float)rho[row * NT + (threadIdx.x & (TPB - 1)) + s * TPB];
  __syncthreads();
  split_row<NT, TPB, RS>(us, sr, RS ? rhos + row * NT : nullptr, k, al, p, t, lds);
  store_row<NT, TPB>(rhs, row, k);
  if (alpha) {
#pragma unroll
    for (int s = 0; s < E; ++s) alpha[row * NT + (threadIdx.x & (TPB - 1)) + s * TPB] = al[s];
  }
}

template <int E, int TPB>
FP_DEV void alpha_half(c32 (&u)[E], const float* rho_row, const fp_nl_params& p,
                       const fp_nl_tables_f32& t) {
  const float hdz = (float)(-0.5 * p.dz);
  const int tid = threadIdx.x & (TPB - 1);
  // alpha of split() (photo-ionisation rate + plasma absorption, NaN -> 0, clipped at 0) for
  // the E samples, with its uniform flags tested once outside the per-sample loops:
  // per-sample branches kept each sample's two table loads in a region of its own, waited
  // for one sample after another
  float a[E];
  if (p.en_pi_loss) {
#pragma unroll
    for (int s = 0; s < E; ++s)
      a[s] = photo_rate(u[s].x * u[s].x + u[s].y * u[s].y, rho_row[tid + s * TPB], p, t);
  } else {
#pragma unroll
    for (int s = 0; s < E; ++s) a[s] = 0.0f;
  }
  if (p.en_plasma_absorb) {
#pragma unroll
    for (int s = 0; s < E; ++s) a[s] = a[s] + (float)p.plasma_pref * rho_row[tid + s * TPB];
  }
#pragma unroll
  for (int s = 0; s < E; ++s) {
    float as = nan0(a[s]);
    as = as > 0.0f ? as : 0.0f;
    u[s] = expf(hdz * as) * u[s];
  }
}

// Fused nonlinear step (alpha/2, RK4, alpha/2) on complex64 rows; spectral in
// and out when the flags are set. cmax_out (optional) receives max |re|, |im|
// of the output rows (float bits, atomicMax) for the next split-FP16 GEMM.
template <int NT> struct StepSmem {
  c32 lds[LdsSize<NT>::value];
  c32 ub[NT];
  float sr[NT];
};

template <int NT, int TPB, bool RS>
FP_DEV void nl_step_row(StepSmem<NT>& sm, size_t row, const float* in, float* out,
                        const double* __restrict__ rho, const double* __restrict__ rhos,
                        int in_spectral, int out_spectral, const fp_nl_params& p,
                        const fp_nl_tables_f32& t, unsigned* __restrict__ cmax_out) {
  constexpr int E = NT / TPB;
  const int tid = threadIdx.x & (TPB - 1);
  c32* lds = sm.lds;
  c32* ub = sm.ub;
  float* sr = sm.sr;
  const c32* tw = reinterpret_cast<const c32*>(t.tw);
  c32 us[E];
  load_row<NT, TPB>(in, row, us);
#pragma unroll
  for (int s = 0; s < E; ++s) sr[tid + s * TPB] = (float)rho[row * NT + tid + s * TPB];
  // The transforms into and out of the row carry the whole field once per z-step: they
  // use the exact variant (residual twiddles tw[NT + m], fused constant rotations), whose
  // round trip has no coherent gain. With the plain float transform each z-step shrank
  // the field by about 1e-7 in amplitude, which grew to a fluence bias of 4e-5 in 200
  // steps. The 16 transforms inside the RK4 stages act on the increments and keep the
  // faster variant.
  if (in_spectral) {
    row_fft<float, NT, TPB, true, true>(us, lds, tw);
    const float invN = 1.0f / (float)NT;
#pragma unroll
    for (int s = 0; s < E; ++s) us[s] = invN * us[s];
  }
  __syncthreads();
  const float* rr = sr;
  const double* rs = RS ? rhos + row * NT : nullptr;
  alpha_half<E, TPB>(us, rr, p, t);
#pragma unroll
  for (int s = 0; s < E; ++s) ub[tid + s * TPB] = us[s];
  const float dz = (float)p.dz;
  c32 acc[E], k[E];
#pragma unroll 1
  for (int stg = 0; stg < 4; ++stg) {
    split_row<NT, TPB, RS>(us, rr, rs, k, nullptr, p, t, lds);
    const float w = (stg == 0 || stg == 3) ? 1.0f : 2.0f;
    const float cdz = (stg < 2) ? 0.5f * dz : dz;
#pragma unroll
    for (int s = 0; s < E; ++s) {
      acc[s] = (stg == 0) ? k[s] : acc[s] + w * k[s];
      us[s] = (stg < 3) ? ub[tid + s * TPB] + cdz * k[s] : ub[tid + s * TPB] + (dz / 6.0f) * acc[s];
    }
  }
  alpha_half<E, TPB>(us, rr, p, t);
  if (out_spectral) row_fft<float, NT, TPB, false, true>(us, lds, tw);
  store_row<NT, TPB>(out, row, us);
  if (cmax_out) {
    float m = 0.0f;
#pragma unroll
    for (int s = 0; s < E; ++s) m = fmaxf(m, fmaxf(fabsf(us[s].x), fabsf(us[s].y)));
    for (int d = 32; d >= 1; d >>= 1) m = fmaxf(m, __shfl_xor(m, d, 64));
    float* sred = reinterpret_cast<float*>(ub);   // the RK base is no longer needed
    __syncthreads();
    if ((tid & 63) == 0) sred[tid >> 6] = m;
    __syncthreads();
    if (tid == 0) {
      for (int w = 1; w < TPB / 64; ++w) m = fmaxf(m, sred[w]);
      atomicMax(cmax_out, __float_as_uint(m));
    }
  }
}

template <int NT, int TPB, bool RS>
__global__ __launch_bounds__(TPB, (TPB >= 512 ? 4 : 2)) void nl_step_kernel(const float* in, float* out,
                                                       const double* __restrict__ rho,
                                                       const double* __restrict__ rhos,
                                                       int in_spectral, int out_spectral,
                                                       fp_nl_params p, fp_nl_tables_f32 t,
                                                       unsigned* __restrict__ cmax_out) {
  __shared__ StepSmem<NT> sm;
  nl_step_row<NT, TPB, RS>(sm, blockIdx.x, in, out, rho, rhos, in_spectral, out_spectral, p, t,
                           cmax_out);
}

}  // namespace fpnl32
