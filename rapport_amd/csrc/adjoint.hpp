// Discrete adjoint of one mixed-precision z-step (complex64 field, FP64 carrier
// scan), one workgroup per radial row. There is no counterpart in the reference
// (D3/sim has no gradients); the forward map being differentiated is the one of
// nl_f32.hpp / plasma.hpp, i.e. D3/sim/integrator.py:138-162.
//
// Convention: for a real loss L of a complex field u, g = dL/dRe u + i dL/dIm u,
// so dL = Re sum conj(g) du; a complex-linear map A has adjoint A^H. Gradients
// between the kernels travel as their FFT along t (gamma = FFT_t g), the same
// representation as the spectral field state, so the Hankel-transform adjoints are
// plain GEMMs with Y^T (Y commutes with the t-FFT).
//
// adj_nl_kernel: for the nonlinear half of the step, v4 = A(RK4(A(v1))) with
//   A(v) = v exp(-alpha(|v|^2, rho) dz/2), recomputes the forward pass in the row
//   (the RK4 stage inputs go to a caller scratch of 4 planes), then applies the
//   adjoints in reverse: A at v3, the four RK4 stages (each the adjoint of the
//   split() right-hand side at that stage input, which also accumulates dL/drho and,
//   with the STE index term i ste_pref rho_s u, dL/drho_s), A at v1.
// adj_plasma_kernel: dL/du from dL/drho through the rate-equation recurrence
//   (kernels.py:36-106, STE off) -- the backward linear recurrence
//   lambda_k = g_k + (d n_{k+1}/d n_k) lambda_{k+1} solved as a parallel scan of
//   affine maps over per-thread chunks -- and adds FFT_t(2 invE2 dL/dI u) to the
//   row of the spectral gradient.
// adj_plasma_ste_kernel: the same with STE on: the state is (n_e, n_s), each step's
//   Jacobian is 2 x 2 (coupled exact-exponential updates, clamps at 0, the
//   n_e + n_s <= n_a rescaling), the chunk maps are 2 x 2 affine maps, and dL/dI also
//   collects the STE rate's LUT derivative.
#pragma once
#include "common.hpp"
#include "fft.hpp"
#include "nl_f32.hpp"
#include "plasma.hpp"
#include "../../include/rapport_amd.h"

namespace fpadj {

using fpfft::LdsSize;
using fpfft::row_fft;
using fpnl32::c32;

// fpnl32::photo_rate with its derivatives d/d|u|^2 and d/drho. The PCHIP piece is
// the same cubic in the local coordinate s = (x - x_i) / h_i, h_i = 10^k / 100.
FP_DEV float photo_rate_d(float a2, float rho, const fp_nl_params& p, const fp_nl_tables_f32& t,
                          float& d_a2, float& d_rho) {
  const float xr = a2 * (float)p.invE2;
  const float x = fminf(fmaxf(xr, (float)p.Imin), (float)p.Imax);
  const int n = p.n_knots;
  float s, ipk;
  const int i = fpnl32::pchip_piece(x, n, s, ipk);
  const float4 c = fpnl32::pchip_coef(t, i);
  const float c0 = c.x, c1 = c.y, c2 = c.z, c3 = c.w;
  const float f = ((c0 * s + c1) * s + c2) * s + c3;
  float dw = ((3.0f * c0 * s + 2.0f * c1) * s + c2) * (100.0f * ipk);
  if (f < 0.0f) dw = -dw;
  if (!(xr > (float)p.Imin && xr < (float)p.Imax)) dw = 0.0f;
  const float irm = fpnl32::rcp((float)p.rho_max);
  const float dl = 1.0f - rho * irm;
  const float depl = fminf(fmaxf(dl, 0.0f), 1.0f);
  const float ddepl = (dl > 0.0f && dl < 1.0f) ? -irm : 0.0f;
  const float iden = fpnl32::rcp(a2 + 1e-30f);
  const float q = nan0(fabsf(f)) * iden;
  d_a2 = nan0(depl * (nan0(dw) * (float)p.invE2 - q) * iden);
  d_rho = q * ddepl;
  return q * depl;
}

// alpha = max(photo + plasma_pref rho, 0) and its derivatives.
FP_DEV float alpha_d(float a2, float rho, const fp_nl_params& p, const fp_nl_tables_f32& t,
                     float& photo, float& p_a2, float& p_rho, float& d_a2, float& d_rho) {
  if (p.en_pi_loss) {
    photo = photo_rate_d(a2, rho, p, t, p_a2, p_rho);
  } else {
    photo = 0.0f; p_a2 = 0.0f; p_rho = 0.0f;
  }
  float a = photo;
  if (p.en_plasma_absorb) a = a + (float)p.plasma_pref * rho;
  a = nan0(a);
  if (a > 0.0f) {
    d_a2 = p_a2;
    d_rho = p_rho + (p.en_plasma_absorb ? (float)p.plasma_pref : 0.0f);
    return a;
  }
  d_a2 = 0.0f;
  d_rho = 0.0f;
  return 0.0f;
}

template <int NT> struct AdjSmem {
  c32 lds[LdsSize<NT>::value];   // FFT exchange
  c32 sG[NT];                    // RK4 base in the forward pass; spectral cotangent in rhs_adj
  c32 g3[NT];                    // gradient at the RK4 output
  float sr[NT];                  // rho row
};

// mu = (d rhs / d y)^H lam at stage input y; grho += (d rhs / d rho)^T lam and, when
// RS (the STE index term is on; rhos_row = the row's rho_s), grhos += (d rhs / d rho_s)^T lam.
template <int NT, int TPB, bool RS>
FP_DEV void rhs_adj_row(const c32 (&y)[NT / TPB], const c32 (&lam)[NT / TPB], c32 (&mu)[NT / TPB],
                        float (&grho)[NT / TPB], float (&grhos)[NT / TPB],
                        const double* __restrict__ rhos_row, AdjSmem<NT>& sm,
                        const fp_nl_params& p, const fp_nl_tables_f32& t) {
  constexpr int E = NT / TPB;
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  const c32* tw = reinterpret_cast<const c32*>(t.tw);
  const c32* Rf = reinterpret_cast<const c32*>(t.Rf);
  const float invN = 1.0f / (float)NT;
  c32 buf[E];
  float gI[E];
#pragma unroll
  for (int s = 0; s < E; ++s) {
    const float a2 = y[s].x * y[s].x + y[s].y * y[s].y;
    buf[s] = (a2 >= 1e-30f) ? lam[s] : mk<float>(0.f, 0.f);
  }
  // alpha u term (pointwise): mu = alpha gp, gI, grho
#pragma unroll
  for (int s = 0; s < E; ++s) {
    const float a2 = y[s].x * y[s].x + y[s].y * y[s].y;
    float ph, pa, pr, da, dr;
    const float al = alpha_d(a2, sm.sr[tid + s * TPB], p, t, ph, pa, pr, da, dr);
    mu[s] = al * buf[s];
    const float re = buf[s].x * y[s].x + buf[s].y * y[s].y;
    gI[s] = da * re;
    grho[s] += dr * re;
  }
  row_fft<float, NT, TPB, false>(buf, sm.lds, tw);
#pragma unroll
  for (int s = 0; s < E; ++s) sm.sG[tid + s * TPB] = buf[s];
  if (p.en_kerr_inst || p.en_kerr_raman) {
    const float fR = (float)p.f_R, kp = (float)p.kerr_pref;
    float K[E];
    if (p.en_kerr_raman) {
      fpraman::raman_row<float, NT, TPB>(y, K, sm.lds, reinterpret_cast<const c32*>(t.rw));
#pragma unroll
      for (int s = 0; s < E; ++s) K[s] = fR * K[s];
    } else {
#pragma unroll
      for (int s = 0; s < E; ++s) K[s] = 0.0f;
    }
    if (p.en_kerr_inst) {
#pragma unroll
      for (int s = 0; s < E; ++s) K[s] += (1.0f - fR) * (y[s].x * y[s].x + y[s].y * y[s].y);
    }
#pragma unroll
    for (int s = 0; s < E; ++s) buf[s] = t.mA[tid + s * TPB] * sm.sG[tid + s * TPB];
    row_fft<float, NT, TPB, true>(buf, sm.lds, tw);
    float gK[E];
#pragma unroll
    for (int s = 0; s < E; ++s) {
      const c32 ga = invN * buf[s];
      const float kk = kp * K[s];
      mu[s] = mu[s] + mk<float>(kk * ga.y, -kk * ga.x);            // (-i kappa K) ga
      gK[s] = -kp * (ga.x * y[s].y - ga.y * y[s].x);               // -kappa Im(conj(ga) y)
      if (p.en_kerr_inst) gI[s] += (1.0f - fR) * gK[s];
    }
    if (p.en_kerr_raman) {
#pragma unroll
      for (int s = 0; s < E; ++s) buf[s] = mk<float>(gK[s], 0.f);
      row_fft<float, NT, TPB, false>(buf, sm.lds, tw);
#pragma unroll
      for (int s = 0; s < E; ++s) buf[s] = buf[s] * conjx(Rf[tid + s * TPB]);
      row_fft<float, NT, TPB, true>(buf, sm.lds, tw);
#pragma unroll
      for (int s = 0; s < E; ++s) gI[s] += fR * buf[s].x * invN;
    }
  }
  if (p.en_pi_loss) {
#pragma unroll
    for (int s = 0; s < E; ++s) buf[s] = t.mB[tid + s * TPB] * sm.sG[tid + s * TPB];
    row_fft<float, NT, TPB, true>(buf, sm.lds, tw);
#pragma unroll
    for (int s = 0; s < E; ++s) {
      const c32 gb = invN * buf[s];
      const float a2 = y[s].x * y[s].x + y[s].y * y[s].y;
      float pa, pr;
      const float ph = photo_rate_d(a2, sm.sr[tid + s * TPB], p, t, pa, pr);
      mu[s] = mu[s] + ph * gb;
      const float re = gb.x * y[s].x + gb.y * y[s].y;
      gI[s] += pa * re;
      grho[s] += pr * re;
    }
  }
  // extra = (q rho + i ste_pref rho_s) y: mu += conj(q rho + i qs) gc, qs = ste_pref rho_s
  // rounded to float as the forward rounds it; d/drho_s = i ste_pref y.
  const bool ste_on = RS && p.ste_pref != 0.0;
  if (p.pc_re != 0.0 || p.pc_im != 0.0 || ste_on) {
    const c32 q = mk<float>((float)(-p.plasma_pref * p.pc_re), (float)(-p.plasma_pref * p.pc_im));
    const float sp = (float)p.ste_pref;
#pragma unroll
    for (int s = 0; s < E; ++s) buf[s] = t.mC[tid + s * TPB] * sm.sG[tid + s * TPB];
    row_fft<float, NT, TPB, true>(buf, sm.lds, tw);
#pragma unroll
    for (int s = 0; s < E; ++s) {
      const c32 gc = invN * buf[s];
      const float r = sm.sr[tid + s * TPB];
      mu[s] = mu[s] + r * (conjx(q) * gc);
      const c32 qy = q * y[s];
      grho[s] += gc.x * qy.x + gc.y * qy.y;
      if (ste_on) {
        const float qs = (float)(p.ste_pref * rhos_row[tid + s * TPB]);
        mu[s] = mu[s] + mk<float>(qs * gc.y, -qs * gc.x);          // (-i qs) gc
        grhos[s] += sp * (gc.y * y[s].x - gc.x * y[s].y);          // Re(conj(gc) i sp y)
      }
    }
  }
#pragma unroll
  for (int s = 0; s < E; ++s) {
    const float f = 2.0f * gI[s];
    mu[s] = mk<float>(nan0(mu[s].x + f * y[s].x), nan0(mu[s].y + f * y[s].y));
  }
}

template <int NT, int TPB>
FP_DEV void load_plane(const float* base, size_t plane, int k, size_t row, c32 (&v)[NT / TPB]) {
  fpnl32::load_row<NT, TPB>(base + (size_t)k * plane, row, v);
}

// Inputs: v1s = FFT_t of the field after the first half linear step (G2 output),
// rho (FP64), rhos (FP64 rho_s, read when RS), g4s = gradient w.r.t. the input of the
// second half linear step (spectral). Outputs: g1s = gradient w.r.t. v1 (spectral),
// grho_out (float), grhos_out (float, written when RS), cmax_out = max |re|, |im| of g1s.
template <int NT, int TPB, bool RS>
__global__ __launch_bounds__(TPB) void adj_nl_kernel(const float* __restrict__ v1s,
                                                     const double* __restrict__ rho,
                                                     const double* __restrict__ rhos,
                                                     const float* __restrict__ g4s,
                                                     float* __restrict__ g1s,
                                                     float* __restrict__ grho_out,
                                                     float* __restrict__ grhos_out,
                                                     float* __restrict__ scratch, fp_nl_params p,
                                                     fp_nl_tables_f32 t,
                                                     unsigned* __restrict__ cmax_out) {
  __shared__ AdjSmem<NT> sm;
  constexpr int E = NT / TPB;
  const size_t row = blockIdx.x;
  const size_t plane = (size_t)p.Nr * NT * 2;       // floats per scratch plane
  const int tid = threadIdx.x;
  const c32* tw = reinterpret_cast<const c32*>(t.tw);
  const float invN = 1.0f / (float)NT;
  const float dz = (float)p.dz, hdz = -0.5f * dz;
  const double* rhos_row = RS ? rhos + row * NT : nullptr;
#pragma unroll
  for (int s = 0; s < E; ++s) sm.sr[tid + s * TPB] = (float)rho[row * NT + tid + s * TPB];
  c32 v[E];
  fpnl32::load_row<NT, TPB>(v1s, row, v);
  row_fft<float, NT, TPB, true>(v, sm.lds, tw);
#pragma unroll
  for (int s = 0; s < E; ++s) v[s] = invN * v[s];
  fpnl32::store_row<NT, TPB>(scratch, row, v);                       // plane 0: v1
  __syncthreads();
  float ph, pa, pr, da, dr;
#pragma unroll
  for (int s = 0; s < E; ++s) {
    const float a2 = v[s].x * v[s].x + v[s].y * v[s].y;
    const float al = alpha_d(a2, sm.sr[tid + s * TPB], p, t, ph, pa, pr, da, dr);
    v[s] = expf(hdz * al) * v[s];
    sm.sG[tid + s * TPB] = v[s];                                    // RK4 base v2
  }
  // forward RK4 (same stage arithmetic as fpnl32::nl_step_row), stage inputs 2-4 to scratch
  {
    c32 acc[E], k[E];
#pragma unroll 1
    for (int stg = 0; stg < 4; ++stg) {
      fpnl32::split_row<NT, TPB, RS>(v, sm.sr, rhos_row, k, nullptr, p, t, sm.lds);
      const float w = (stg == 0 || stg == 3) ? 1.0f : 2.0f;
      const float cdz = (stg < 2) ? 0.5f * dz : dz;
#pragma unroll
      for (int s = 0; s < E; ++s) {
        acc[s] = (stg == 0) ? k[s] : acc[s] + w * k[s];
        v[s] = (stg < 3) ? sm.sG[tid + s * TPB] + cdz * k[s]
                         : sm.sG[tid + s * TPB] + (dz / 6.0f) * acc[s];
      }
      if (stg < 3) fpnl32::store_row<NT, TPB>(scratch + (size_t)(stg + 1) * plane, row, v);
    }
  }
  // v = v3. Gradient w.r.t. v4 in the time domain.
  c32 g[E];
  fpnl32::load_row<NT, TPB>(g4s, row, g);
  row_fft<float, NT, TPB, true>(g, sm.lds, tw);
  float grho[E], grhos[E];
#pragma unroll
  for (int s = 0; s < E; ++s) {
    grhos[s] = 0.0f;
    g[s] = invN * g[s];
    const float a2 = v[s].x * v[s].x + v[s].y * v[s].y;
    const float al = alpha_d(a2, sm.sr[tid + s * TPB], p, t, ph, pa, pr, da, dr);
    const float e = expf(hdz * al);
    const float sc = (g[s].x * v[s].x + g[s].y * v[s].y) * e * hdz;
    g[s] = mk<float>(e * g[s].x + 2.0f * sc * da * v[s].x, e * g[s].y + 2.0f * sc * da * v[s].y);
    grho[s] = sc * dr;
    sm.g3[tid + s * TPB] = g[s];
  }
  // RK4 adjoint: lambda_k4 = dz/6 g3, lambda_k3 = dz/3 g3 + dz mu4,
  // lambda_k2 = dz/3 g3 + dz/2 mu3, lambda_k1 = dz/6 g3 + dz/2 mu2; g2 = g3 + sum mu.
  c32 mu[E], lam[E], y[E];
#pragma unroll
  for (int s = 0; s < E; ++s) {
    v[s] = mk<float>(0.f, 0.f);                  // v: running sum of mu
    mu[s] = mk<float>(0.f, 0.f);
  }
#pragma unroll 1
  for (int stg = 3; stg >= 0; --stg) {
    const float cg = (stg == 0 || stg == 3) ? dz / 6.0f : dz / 3.0f;
    const float cm = (stg == 3) ? 0.0f : (stg == 2 ? dz : 0.5f * dz);
#pragma unroll
    for (int s = 0; s < E; ++s) lam[s] = cg * sm.g3[tid + s * TPB] + cm * mu[s];
    if (stg > 0) {
      load_plane<NT, TPB>(scratch, plane, stg, row, y);
    } else {
      load_plane<NT, TPB>(scratch, plane, 0, row, y);
#pragma unroll
      for (int s = 0; s < E; ++s) {
        const float a2 = y[s].x * y[s].x + y[s].y * y[s].y;
        const float al = alpha_d(a2, sm.sr[tid + s * TPB], p, t, ph, pa, pr, da, dr);
        y[s] = expf(hdz * al) * y[s];
      }
    }
    rhs_adj_row<NT, TPB, RS>(y, lam, mu, grho, grhos, rhos_row, sm, p, t);
#pragma unroll
    for (int s = 0; s < E; ++s) v[s] = v[s] + mu[s];
  }
  // g2 = g3 + sum mu; adjoint of A at v1
  load_plane<NT, TPB>(scratch, plane, 0, row, y);
#pragma unroll
  for (int s = 0; s < E; ++s) {
    const c32 g2 = sm.g3[tid + s * TPB] + v[s];
    const float a2 = y[s].x * y[s].x + y[s].y * y[s].y;
    const float al = alpha_d(a2, sm.sr[tid + s * TPB], p, t, ph, pa, pr, da, dr);
    const float e = expf(hdz * al);
    const float sc = (g2.x * y[s].x + g2.y * y[s].y) * e * hdz;
    g[s] = mk<float>(e * g2.x + 2.0f * sc * da * y[s].x, e * g2.y + 2.0f * sc * da * y[s].y);
    grho[s] += sc * dr;
    grho_out[row * NT + tid + s * TPB] = nan0(grho[s]);
    if (RS) grhos_out[row * NT + tid + s * TPB] = nan0(grhos[s]);
  }
  row_fft<float, NT, TPB, false>(g, sm.lds, tw);
  fpnl32::store_row<NT, TPB>(g1s, row, g);
  if (cmax_out) {
    float m = 0.0f;
#pragma unroll
    for (int s = 0; s < E; ++s) m = fmaxf(m, fmaxf(fabsf(g[s].x), fabsf(g[s].y)));
    for (int d = 32; d >= 1; d >>= 1) m = fmaxf(m, __shfl_xor(m, d, 64));
    float* sred = sm.sr;
    __syncthreads();
    if ((tid & 63) == 0) sred[tid >> 6] = m;
    __syncthreads();
    if (tid == 0) {
      for (int w = 1; w < TPB / 64; ++w) m = fmaxf(m, sred[w]);
      atomicMax(cmax_out, __float_as_uint(m));
    }
  }
}

// The log-LUT rate of fpplasma::lut_rate and its derivative dW/dI.
FP_DEV double lut_rate_d(double I, const double* __restrict__ tab, const fp_plasma_params& p,
                         double& dW) {
  double Ic = (!fp_isfinite(I) || I < 0.0) ? 0.0 : I;
  const bool inside = Ic > p.Imin && Ic < p.Imax;
  if (Ic < p.Imin) Ic = p.Imin;
  if (Ic > p.Imax) Ic = p.Imax;
  const double fidx = (log10(Ic) - p.log_min_I) * p.inv_log_step;
  int idx = (int)fidx;
  if (idx < 0) idx = 0;
  if (idx > p.table_size - 2) idx = p.table_size - 2;
  const double r = tab[idx] + (fidx - (double)idx) * (tab[idx + 1] - tab[idx]);
  const bool ok = fp_isfinite(r) && r >= 0.0;
  dW = (inside && ok) ? (tab[idx + 1] - tab[idx]) * p.inv_log_step / (Ic * 2.302585092994046) : 0.0;
  return ok ? r : 0.0;
}

template <int NT, int TPB>
struct AdjPlasmaSmem {
  double sI[NT];
  double sGI[NT];
  double sA[TPB], sB[TPB];
  c32 lds[LdsSize<NT>::value];
};

// us: spectral field row u_n (the state the forward scan read); rho: the forward
// densities of that state; grho: dL/drho; gs: spectral gradient row, += FFT_t of
// 2 invE2 (dL/dI) u.
template <int NT, int TPB>
__global__ __launch_bounds__(TPB) void adj_plasma_kernel(const float* __restrict__ us,
                                                         const double* __restrict__ rho,
                                                         const float* __restrict__ grho,
                                                         float* __restrict__ gs,
                                                         fp_plasma_params p,
                                                         const double* __restrict__ Wg,
                                                         const float* __restrict__ tw_,
                                                         unsigned* __restrict__ cmax_out) {
  constexpr int E = NT / TPB;      // strided FFT layout and contiguous chunk length
  __shared__ AdjPlasmaSmem<NT, TPB> sm;
  const size_t row = blockIdx.x;
  const int tid = threadIdx.x;
  const c32* tw = reinterpret_cast<const c32*>(tw_);
  c32 u[E];
  fpnl32::load_row<NT, TPB>(us, row, u);
  row_fft<float, NT, TPB, true>(u, sm.lds, tw);
  const float invN = 1.0f / (float)NT;
#pragma unroll
  for (int s = 0; s < E; ++s) {
    u[s] = invN * u[s];
    const double x = (double)u[s].x, y = (double)u[s].y;
    sm.sI[tid + s * TPB] = (x * x + y * y) * p.field_to_I;
  }
  __syncthreads();
  const int k0 = tid * E;
  const double* nrow = rho + row * NT;
  const float* grow = grho + row * NT;
  const double na = p.na, dt = p.dt, beta = p.beta_g, itr = p.inv_tau_r;
  double a[E], rp[E], rm[E];
  double dWc;
  double Wc = lut_rate_d(sm.sI[k0], Wg, p, dWc);
#pragma unroll
  for (int j = 0; j < E; ++j) {
    const int k = k0 + j;
    if (k == NT - 1) {
      a[j] = 0.0; rp[j] = 0.0; rm[j] = 0.0;
      continue;
    }
    const double Ic = sm.sI[k], In = sm.sI[k + 1];
    double dWn;
    const double Wn = lut_rate_d(In, Wg, p, dWn);
    const double n = nrow[k];
    const double Ia = 0.5 * (Ic + In), Wa = 0.5 * (Wc + Wn);
    const double dl = 1.0 - n / na;
    const double depl = dl < 0.0 ? 0.0 : (dl > 1.0 ? 1.0 : dl);
    const double dd = (dl > 0.0 && dl < 1.0) ? -1.0 / na : 0.0;
    const double S = Wa * depl, L = beta * Ia * depl - itr;
    const fpplasma::Step st = fpplasma::exp_step(n, S, L, dt);
    const bool ok = fp_isfinite(st.xnew) && st.xnew >= 0.0 && st.xnew <= na;
    const double dxdL = dt * st.E * n + S * dt * dt * st.dphi1;
    const double dxdS = dt * st.phi1;
    a[j] = ok ? st.E + dxdS * Wa * dd + dxdL * beta * Ia * dd : 0.0;
    const double pk = ok ? dxdL * beta * depl : 0.0;
    const double qk = ok ? dxdS * depl : 0.0;
    rp[j] = pk + qk * dWc;
    rm[j] = pk + qk * dWn;
    Wc = Wn;
    dWc = dWn;
  }
  // chunk map lambda_{k0} = A lambda_{k0 + E} + B
  double A = 1.0, B = 0.0;
#pragma unroll
  for (int j = E - 1; j >= 0; --j) {
    B = a[j] * B + (double)grow[k0 + j];
    A = a[j] * A;
  }
  sm.sA[tid] = A;
  sm.sB[tid] = B;
  __syncthreads();
  for (int off = 1; off < TPB; off <<= 1) {   // suffix composition M_t o M_{t+1} o ...
    double A2 = 1.0, B2 = 0.0;
    const bool has = tid + off < TPB;
    if (has) { A2 = sm.sA[tid + off]; B2 = sm.sB[tid + off]; }
    __syncthreads();
    if (has) { B = A * B2 + B; A = A * A2; }
    sm.sA[tid] = A;
    sm.sB[tid] = B;
    __syncthreads();
  }
  double lam_next = (tid + 1 < TPB) ? sm.sB[tid + 1] : 0.0;   // lambda_{k0 + E}
  const double carry = 0.5 * lam_next * rm[E - 1];             // -> dL/dI at k0 + E
  double gown[E];
#pragma unroll
  for (int j = E - 1; j >= 0; --j) {
    gown[j] = 0.5 * lam_next * rp[j];
    const double lam = a[j] * lam_next + (double)grow[k0 + j];
    if (j > 0) gown[j] += 0.5 * lam * rm[j - 1];
    lam_next = lam;
  }
  __syncthreads();          // every thread is done reading sA / sB
#pragma unroll
  for (int j = 0; j < E; ++j) sm.sGI[k0 + j] = gown[j];
  __syncthreads();
  if (k0 + E < NT) sm.sGI[k0 + E] += carry;
  __syncthreads();
  const double f2 = 2.0 * p.field_to_I;
#pragma unroll
  for (int s = 0; s < E; ++s) {
    const float c = (float)(f2 * sm.sGI[tid + s * TPB]);
    u[s] = c * u[s];
  }
  row_fft<float, NT, TPB, false>(u, sm.lds, tw);
  c32* grow_s = reinterpret_cast<c32*>(gs) + row * NT;
  float m = 0.0f;
#pragma unroll
  for (int s = 0; s < E; ++s) {
    const c32 r = grow_s[tid + s * TPB] + u[s];
    grow_s[tid + s * TPB] = r;
    m = fmaxf(m, fmaxf(fabsf(r.x), fabsf(r.y)));
  }
  if (cmax_out) {
    for (int d = 32; d >= 1; d >>= 1) m = fmaxf(m, __shfl_xor(m, d, 64));
    __syncthreads();
    float* sred = reinterpret_cast<float*>(sm.sA);
    if ((tid & 63) == 0) sred[tid >> 6] = m;
    __syncthreads();
    if (tid == 0) {
      for (int w = 1; w < TPB / 64; ++w) m = fmaxf(m, sred[w]);
      atomicMax(cmax_out, __float_as_uint(m));
    }
  }
}

// One step of the carrier recurrence with STE (fpplasma::rate_step<true>) linearised
// for the backward sweep: lambda_k = g_k + M lambda_{k+1}, lambda = dL/d(n_e, n_s),
// and dL/dI_avg = ie le + is ls, dL/dW_avg = we le + ws ls, dL/dWs_avg = se le + ss ls
// for lambda_{k+1} = (le, ls). a = E(n_e; S_e, L_e) and b = E(n_s; S_s, L_s) are the
// clamped exact-exponential updates; where a + b > n_a the step returns (a, b) n_a /
// (a + b), whose Jacobian n_a / (a + b)^2 [[b, -a], [-b, a]] is applied first.
struct StepAdj {
  double m00, m01, m10, m11;
  double ie, is, we, ws, se, ss;
};

FP_DEV StepAdj ste_step_adj(double ne, double ns, double Ia, double Wa, double Wsa,
                            const fp_plasma_params& p) {
  const double na = p.na, dt = p.dt, bg = p.beta_g, bs = p.beta_s, itr = p.inv_tau_r;
  const double dl = 1.0 - ne / na;
  const double depl = dl < 0.0 ? 0.0 : (dl > 1.0 ? 1.0 : dl);
  const double dd = (dl > 0.0 && dl < 1.0) ? -1.0 / na : 0.0;
  const double wst = Wsa + bs * Ia * ne;
  const double Se = Wa * depl + wst * (ns / na);
  const double Le = bg * Ia * depl - itr;
  const fpplasma::Step sa = fpplasma::exp_step(ne, Se, Le, dt);
  const bool oka = fp_isfinite(sa.xnew) && sa.xnew >= 0.0;
  const double a = oka ? sa.xnew : 0.0;
  const double aS = dt * sa.phi1, aL = dt * sa.E * ne + Se * dt * dt * sa.dphi1;
  const double a_ne = sa.E + aS * (Wa * dd + bs * Ia * (ns / na)) + aL * bg * Ia * dd;
  const double a_ns = aS * wst / na;
  const double a_I = aS * bs * ne * (ns / na) + aL * bg * depl;
  const double a_W = aS * depl;
  const double a_Ws = aS * (ns / na);
  const double Ss = itr * ne, Ls = -wst / na - p.inv_tau_ste;
  const fpplasma::Step sb = fpplasma::exp_step(ns, Ss, Ls, dt);
  const bool okb = fp_isfinite(sb.xnew) && sb.xnew >= 0.0;
  const double b = okb ? sb.xnew : 0.0;
  const double bS = dt * sb.phi1, bL = dt * sb.E * ns + Ss * dt * dt * sb.dphi1;
  const double b_ne = bS * itr - bL * bs * Ia / na;
  const double b_ns = sb.E;
  const double b_I = -bL * bs * ne / na;
  const double b_Ws = -bL / na;
  // mu_a = rae le + ras ls, mu_b = rbe le + rbs ls: dL/d(a, b)
  double rae = 1.0, ras = 0.0, rbe = 0.0, rbs = 1.0;
  const double tot = a + b;
  if (tot > na) {
    const double f = na / (tot * tot);
    rae = f * b; ras = -f * b; rbe = -f * a; rbs = f * a;
  }
  if (!oka) { rae = 0.0; ras = 0.0; }
  if (!okb) { rbe = 0.0; rbs = 0.0; }
  StepAdj s;
  s.m00 = rae * a_ne + rbe * b_ne;
  s.m01 = ras * a_ne + rbs * b_ne;
  s.m10 = rae * a_ns + rbe * b_ns;
  s.m11 = ras * a_ns + rbs * b_ns;
  s.ie = rae * a_I + rbe * b_I;
  s.is = ras * a_I + rbs * b_I;
  s.we = rae * a_W;
  s.ws = ras * a_W;
  s.se = rae * a_Ws + rbe * b_Ws;
  s.ss = ras * a_Ws + rbs * b_Ws;
  return s;
}

template <int NT, int TPB>
struct AdjPlasmaSteSmem {
  double sI[NT];
  double sGI[NT];
  double sM[6][TPB];     // chunk maps (A row-major, B) for the suffix scan
  c32 lds[LdsSize<NT>::value];
};

// us, rho, rhos: the spectral state the forward scan read and its densities; grho,
// grhos: dL/drho, dL/drho_s; gs: spectral gradient row, += FFT_t of 2 invE2 (dL/dI) u.
// Each thread owns NT/TPB consecutive samples. Pass 1 composes its chunk's affine map
// lambda_{k0} = A lambda_{k0+E} + B backwards, the suffix scan over threads gives each
// chunk its incoming lambda_{k0+E}, and pass 2 re-linearises every step (the LUT rates
// and step Jacobians are recomputed rather than kept: 10 doubles a step would not fit
// in registers) and accumulates dL/dI in LDS.
template <int NT, int TPB>
__global__ __launch_bounds__(TPB) void adj_plasma_ste_kernel(
    const float* __restrict__ us, const double* __restrict__ rho, const double* __restrict__ rhos,
    const float* __restrict__ grho, const float* __restrict__ grhos, float* __restrict__ gs,
    fp_plasma_params p, const double* __restrict__ Wg, const double* __restrict__ Wsg,
    const float* __restrict__ tw_, unsigned* __restrict__ cmax_out) {
  constexpr int E = NT / TPB;
  __shared__ AdjPlasmaSteSmem<NT, TPB> sm;
  const size_t row = blockIdx.x;
  const int tid = threadIdx.x;
  const c32* tw = reinterpret_cast<const c32*>(tw_);
  c32 u[E];
  fpnl32::load_row<NT, TPB>(us, row, u);
  row_fft<float, NT, TPB, true>(u, sm.lds, tw);
  const float invN = 1.0f / (float)NT;
#pragma unroll
  for (int s = 0; s < E; ++s) {
    u[s] = invN * u[s];
    const double x = (double)u[s].x, y = (double)u[s].y;
    sm.sI[tid + s * TPB] = (x * x + y * y) * p.field_to_I;
  }
  __syncthreads();
  const int k0 = tid * E;
  const double* nrow = rho + row * NT;
  const double* srow = rhos + row * NT;
  const float* ge = grho + row * NT;
  const float* gsr = grhos + row * NT;
  // rates at the chunk's right neighbour k0 + E (the next chunk's first sample)
  double Wr = 0.0, dWr = 0.0, Wsr = 0.0, dWsr = 0.0;
  if (k0 + E < NT) {
    Wr = lut_rate_d(sm.sI[k0 + E], Wg, p, dWr);
    Wsr = lut_rate_d(sm.sI[k0 + E], Wsg, p, dWsr);
  }
  // pass 1: chunk map
  double A00 = 1.0, A01 = 0.0, A10 = 0.0, A11 = 1.0, B0 = 0.0, B1 = 0.0;
  {
    double W1 = Wr, Ws1 = Wsr;
#pragma unroll 1
    for (int j = E - 1; j >= 0; --j) {
      const int k = k0 + j;
      double d0;
      const double W0 = lut_rate_d(sm.sI[k], Wg, p, d0);
      const double Ws0 = lut_rate_d(sm.sI[k], Wsg, p, d0);
      if (k < NT - 1) {
        const StepAdj st = ste_step_adj(nrow[k], srow[k], 0.5 * (sm.sI[k] + sm.sI[k + 1]),
                                        0.5 * (W0 + W1), 0.5 * (Ws0 + Ws1), p);
        const double n00 = st.m00 * A00 + st.m01 * A10, n01 = st.m00 * A01 + st.m01 * A11;
        const double n10 = st.m10 * A00 + st.m11 * A10, n11 = st.m10 * A01 + st.m11 * A11;
        const double c0 = st.m00 * B0 + st.m01 * B1 + (double)ge[k];
        const double c1 = st.m10 * B0 + st.m11 * B1 + (double)gsr[k];
        A00 = n00; A01 = n01; A10 = n10; A11 = n11; B0 = c0; B1 = c1;
      } else {                                   // last sample: lambda_{NT-1} = g_{NT-1}
        A00 = A01 = A10 = A11 = 0.0;
        B0 = (double)ge[k];
        B1 = (double)gsr[k];
      }
      W1 = W0;
      Ws1 = Ws0;
    }
  }
  // suffix composition F_t o F_{t+1} o ... : A = A_t A', B = A_t B' + B_t
  sm.sM[0][tid] = A00; sm.sM[1][tid] = A01; sm.sM[2][tid] = A10; sm.sM[3][tid] = A11;
  sm.sM[4][tid] = B0; sm.sM[5][tid] = B1;
  __syncthreads();
  for (int off = 1; off < TPB; off <<= 1) {
    const bool has = tid + off < TPB;
    double o00 = 1.0, o01 = 0.0, o10 = 0.0, o11 = 1.0, ob0 = 0.0, ob1 = 0.0;
    if (has) {
      o00 = sm.sM[0][tid + off]; o01 = sm.sM[1][tid + off];
      o10 = sm.sM[2][tid + off]; o11 = sm.sM[3][tid + off];
      ob0 = sm.sM[4][tid + off]; ob1 = sm.sM[5][tid + off];
    }
    __syncthreads();
    if (has) {
      const double c0 = A00 * ob0 + A01 * ob1 + B0, c1 = A10 * ob0 + A11 * ob1 + B1;
      const double n00 = A00 * o00 + A01 * o10, n01 = A00 * o01 + A01 * o11;
      const double n10 = A10 * o00 + A11 * o10, n11 = A10 * o01 + A11 * o11;
      A00 = n00; A01 = n01; A10 = n10; A11 = n11; B0 = c0; B1 = c1;
    }
    sm.sM[0][tid] = A00; sm.sM[1][tid] = A01; sm.sM[2][tid] = A10; sm.sM[3][tid] = A11;
    sm.sM[4][tid] = B0; sm.sM[5][tid] = B1;
    __syncthreads();
  }
  // lambda_{k0+E}: the suffix of the next chunk applied to lambda_NT = 0
  double le = (tid + 1 < TPB) ? sm.sM[4][tid + 1] : 0.0;
  double ls = (tid + 1 < TPB) ? sm.sM[5][tid + 1] : 0.0;
  // pass 2: dL/dI of the chunk's samples (sGI[k0 .. k0+E-1] are this thread's)
#pragma unroll 1
  for (int j = 0; j < E; ++j) sm.sGI[k0 + j] = 0.0;
  double carry = 0.0;                            // -> dL/dI at k0 + E
  {
    double W1 = Wr, dW1 = dWr, Ws1 = Wsr, dWs1 = dWsr;
#pragma unroll 1
    for (int j = E - 1; j >= 0; --j) {
      const int k = k0 + j;
      double dW0, dWs0;
      const double W0 = lut_rate_d(sm.sI[k], Wg, p, dW0);
      const double Ws0 = lut_rate_d(sm.sI[k], Wsg, p, dWs0);
      if (k < NT - 1) {
        const StepAdj st = ste_step_adj(nrow[k], srow[k], 0.5 * (sm.sI[k] + sm.sI[k + 1]),
                                        0.5 * (W0 + W1), 0.5 * (Ws0 + Ws1), p);
        const double gIa = st.ie * le + st.is * ls;
        const double gWa = st.we * le + st.ws * ls;
        const double gWsa = st.se * le + st.ss * ls;
        sm.sGI[k] += 0.5 * (gIa + gWa * dW0 + gWsa * dWs0);
        const double right = 0.5 * (gIa + gWa * dW1 + gWsa * dWs1);
        if (j + 1 < E) sm.sGI[k + 1] += right;
        else carry = right;
        const double nle = (double)ge[k] + st.m00 * le + st.m01 * ls;
        const double nls = (double)gsr[k] + st.m10 * le + st.m11 * ls;
        le = nle;
        ls = nls;
      } else {
        le = (double)ge[k];
        ls = (double)gsr[k];
      }
      W1 = W0; dW1 = dW0; Ws1 = Ws0; dWs1 = dWs0;
    }
  }
  __syncthreads();
  if (k0 + E < NT) sm.sGI[k0 + E] += carry;
  __syncthreads();
  const double f2 = 2.0 * p.field_to_I;
#pragma unroll
  for (int s = 0; s < E; ++s) {
    const float c = (float)(f2 * sm.sGI[tid + s * TPB]);
    u[s] = c * u[s];
  }
  row_fft<float, NT, TPB, false>(u, sm.lds, tw);
  c32* grow_s = reinterpret_cast<c32*>(gs) + row * NT;
  float m = 0.0f;
#pragma unroll
  for (int s = 0; s < E; ++s) {
    const c32 r = grow_s[tid + s * TPB] + u[s];
    grow_s[tid + s * TPB] = r;
    m = fmaxf(m, fmaxf(fabsf(r.x), fabsf(r.y)));
  }
  if (cmax_out) {
    for (int d = 32; d >= 1; d >>= 1) m = fmaxf(m, __shfl_xor(m, d, 64));
    __syncthreads();
    float* sred = reinterpret_cast<float*>(&sm.sM[0][0]);
    if ((tid & 63) == 0) sred[tid >> 6] = m;
    __syncthreads();
    if (tid == 0) {
      for (int w = 1; w < TPB / 64; ++w) m = fmaxf(m, sred[w]);
      atomicMax(cmax_out, __float_as_uint(m));
    }
  }
}

}  // namespace fpadj
