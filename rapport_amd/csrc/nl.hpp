// Nonlinear operator of the split step, one workgroup per radial row.
//
// split() (D3/sim/operators.py:89-146) evaluated on a row held in registers in
// the FFT stride layout (fft.hpp): pointwise Keldysh/plasma terms, the Raman
// convolution and the three spectrally weighted transforms all happen in LDS
// without touching HBM. The fused step kernel chains the alpha half-steps and
// the four RK4 stages of Integrator.step (D3/sim/integrator.py:141-147) on the
// same row, so the whole nonlinear part of a z-step reads the row once and
// writes it once.
#pragma once
#include "common.hpp"
#include "fft.hpp"
#include "raman.hpp"
#include "../../include/rapport_amd.h"

namespace fpnl {

using fpfft::LdsSize;
using fpfft::row_fft;

// Exact decade boundaries of the PCHIP knot grid 10^k (1 + 0.01 j).
__device__ const double kPow10[18] = {1e0, 1e1, 1e2, 1e3, 1e4, 1e5, 1e6, 1e7, 1e8,
                                      1e9, 1e10, 1e11, 1e12, 1e13, 1e14, 1e15, 1e16, 1e17};
__device__ const double kInvPow10[18] = {1e0, 1e-1, 1e-2, 1e-3, 1e-4, 1e-5, 1e-6, 1e-7, 1e-8,
                                         1e-9, 1e-10, 1e-11, 1e-12, 1e-13, 1e-14, 1e-15, 1e-16,
                                         1e-17};

// Interval i with knots[i] <= x < knots[i+1] (x in [knots[0], knots[n-1]];
// x == knots[n-1] maps to n-2, as scipy.interpolate.PPoly does), and xi = knots[i].
// Branch-free, so the table loads of a thread's samples overlap: the decade follows from
// the binary exponent (floor(e2 log10 2) = (e2 * 1233) >> 12, floor(log10 x) or one below
// it, settled by one comparison with 10^(k+1)); the index from the scaled mantissa is then
// at most one interval off, and one comparison each way with the knots makes it exact
// (checked on the CPU against searchsorted on every knot, its two neighbouring doubles and
// 2e6 random intensities). The loop form waited for each load in turn.
FP_DEV int pchip_interval(double x, const double* __restrict__ knots, int n, double& xi) {
  const int e2 = (int)((__double_as_longlong(x) >> 52) & 0x7ff) - 1023;
  int k = e2 <= 0 ? 0 : (e2 * 1233) >> 12;
  if (k > 16) k = 16;
  const double p1 = kPow10[k + 1], ip0 = kInvPow10[k], ip1 = kInvPow10[k + 1];
  const bool up = k < 16 && x >= p1;
  k += up ? 1 : 0;
  int j = (int)((x * (up ? ip1 : ip0) - 1.0) * 100.0);
  j = j < 0 ? 0 : (j > 899 ? 899 : j);
  int i = k * 900 + j;
  i = i > n - 2 ? n - 2 : i;
  const double km = knots[i > 0 ? i - 1 : 0], k0 = knots[i], k1 = knots[i + 1];
  const bool inc = i < n - 2 && k1 <= x;
  const bool dec = i > 0 && k0 > x;
  xi = inc ? k1 : (dec ? km : k0);
  return inc ? i + 1 : (dec ? i - 1 : i);
}

// |PCHIP(I)| with nan_to_num, FP64 power basis in dx = x - x_i.
FP_DEV double keldysh_W(double absu2, const fp_nl_params& p, const fp_nl_tables& t) {
  double x = absu2 * p.invE2;
  x = fmin(fmax(x, p.Imin), p.Imax);   // cp.clip
  if (!fp_isfinite(x)) x = p.Imax;     // clip of +inf; NaN handled below
  const int n = p.n_knots;
  double xi;
  const int i = pchip_interval(x, t.knots, n, xi);
  const double dx = x - xi;
  const int m = n - 1;
  const double c0 = t.pcoef[i], c1 = t.pcoef[m + i], c2 = t.pcoef[2 * m + i], c3 = t.pcoef[3 * m + i];
  const double w = fabs(((c0 * dx + c1) * dx + c2) * dx + c3);
  return nan0(w);
}

// photo and alpha of split() (operators.py:91-107) for one sample.
FP_DEV void photo_alpha(double absu2, double rho, const fp_nl_params& p, const fp_nl_tables& t,
                        double& photo, double& alpha) {
  double ph = 0.0;
  if (p.en_pi_loss) {
    const double W = keldysh_W(absu2, p, t);
    double depl = 1.0 - rho / p.rho_max;
    depl = fmin(fmax(depl, 0.0), 1.0);
    ph = W * p.photo_pref / (absu2 + 1e-30) * depl;
  }
  double a = ph;
  if (p.en_plasma_absorb) a = a + p.plasma_pref * rho;
  a = nan0(a);
  photo = ph;
  alpha = a > 0.0 ? a : 0.0;
}

// One split() evaluation on a row in registers: k = rhs(us) and, if al is
// given, al = alpha(us). Register state is kept to us and two transform
// buffers: rho (and rho_s) are re-read from the row in global memory (L2
// resident), |us|^2 is recomputed where needed, the Kerr intensity is formed in
// the transform buffer, and the photo-ionisation rate of each sample is parked in the
// LDS row sph (NT doubles) for the final alpha*u term, which is alpha of the same field.
template <int NT, int TPB, bool RS>
FP_DEV void split_row(const cpx<double> (&us)[NT / TPB], const double* __restrict__ rho_row,
                      const double* __restrict__ rhos_row, cpx<double> (&k)[NT / TPB],
                      double* al, const fp_nl_params& p, const fp_nl_tables& t,
                      cpx<double>* lds, double* sph) {
  constexpr int E = NT / TPB;
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));   // recompute table addresses here rather than hoist them
  const cpx<double>* tw = reinterpret_cast<const cpx<double>*>(t.tw);
  const cpx<double>* Rf = reinterpret_cast<const cpx<double>*>(t.Rf);
  const double invN = 1.0 / (double)NT;
  cpx<double> buf[E];
  // k doubles as the spectral accumulator of the three weighted transforms.
  // B = photo * u
  if (p.en_pi_loss) {
#pragma unroll
    for (int s = 0; s < E; ++s) {
      const double a2 = us[s].x * us[s].x + us[s].y * us[s].y;
      double ph, a;
      photo_alpha(a2, rho_row[tid + s * TPB], p, t, ph, a);
      sph[tid + s * TPB] = ph;
      buf[s] = ph * us[s];
    }
    row_fft<double, NT, TPB, false>(buf, lds, tw);
#pragma unroll
    for (int s = 0; s < E; ++s) k[s] = t.mB[tid + s * TPB] * buf[s];
  } else {
#pragma unroll
    for (int s = 0; s < E; ++s) k[s] = mk<double>(0.0, 0.0);
  }
  // Kerr intensity (instantaneous + Raman convolution), then A = i kappa kerr_I u
  if (p.en_kerr_inst || p.en_kerr_raman) {
    double KR[E];
    if (p.en_kerr_raman)
      fpraman::raman_row<double, NT, TPB>(us, KR, lds, reinterpret_cast<const cpx<double>*>(t.rw));
#pragma unroll
    for (int s = 0; s < E; ++s) {
      const double a2 = us[s].x * us[s].x + us[s].y * us[s].y;
      double kerr = p.en_kerr_inst ? (1.0 - p.f_R) * a2 : 0.0;
      if (p.en_kerr_raman) kerr = kerr + p.f_R * KR[s];
      const double kk = p.kerr_pref * kerr;
      buf[s] = mk<double>(-kk * us[s].y, kk * us[s].x);
    }
    row_fft<double, NT, TPB, false>(buf, lds, tw);
#pragma unroll
    for (int s = 0; s < E; ++s) k[s] = k[s] + t.mA[tid + s * TPB] * buf[s];
  }
  // extra = -pp * pc * rho * u + i * ste * rho_s * u
  const bool any_extra = (p.pc_re != 0.0 || p.pc_im != 0.0) || (RS && p.ste_pref != 0.0);
  if (any_extra) {
    const cpx<double> g = mk<double>(-p.plasma_pref * p.pc_re, -p.plasma_pref * p.pc_im);
#pragma unroll
    for (int s = 0; s < E; ++s) {
      cpx<double> e = rho_row[tid + s * TPB] * (g * us[s]);
      if (RS && p.ste_pref != 0.0) {
        const double q = p.ste_pref * rhos_row[tid + s * TPB];
        e = e + mk<double>(-q * us[s].y, q * us[s].x);
      }
      buf[s] = e;
    }
    row_fft<double, NT, TPB, false>(buf, lds, tw);
#pragma unroll
    for (int s = 0; s < E; ++s) k[s] = k[s] + t.mC[tid + s * TPB] * buf[s];
  }
  row_fft<double, NT, TPB, true>(k, lds, tw);
#pragma unroll
  for (int s = 0; s < E; ++s) {
    const double a2 = us[s].x * us[s].x + us[s].y * us[s].y;
    // alpha as photo_alpha forms it, from the rate computed above for this field
    double a = p.en_pi_loss ? sph[tid + s * TPB] : 0.0;
    if (p.en_plasma_absorb) a = a + p.plasma_pref * rho_row[tid + s * TPB];
    a = nan0(a);
    a = a > 0.0 ? a : 0.0;
    if (al) al[s] = a;
    cpx<double> r = mk<double>(k[s].x * invN + a * us[s].x, k[s].y * invN + a * us[s].y);
    if (!(a2 >= 1e-30)) r = mk<double>(0.0, 0.0);   // where(absu2 < 1e-30, 0, rhs)
    k[s] = mk<double>(nan0(r.x), nan0(r.y));
  }
}

template <int NT, int TPB>
FP_DEV void load_row(const double* __restrict__ src, size_t row, cpx<double> (&v)[NT / TPB]) {
  const cpx<double>* r = reinterpret_cast<const cpx<double>*>(src) + row * NT;
#pragma unroll
  for (int s = 0; s < NT / TPB; ++s) v[s] = r[threadIdx.x + s * TPB];
}
template <int NT, int TPB>
FP_DEV void store_row(double* __restrict__ dst, size_t row, const cpx<double> (&v)[NT / TPB]) {
  cpx<double>* r = reinterpret_cast<cpx<double>*>(dst) + row * NT;
#pragma unroll
  for (int s = 0; s < NT / TPB; ++s) r[threadIdx.x + s * TPB] = v[s];
}

// split(): rhs and alpha of a time-domain field.
template <int NT, int TPB, bool RS>
__global__ __launch_bounds__(TPB) void split_kernel(const double* __restrict__ u,
                                                     const double* __restrict__ rho,
                                                     const double* __restrict__ rhos,
                                                     double* __restrict__ rhs,
                                                     double* __restrict__ alpha, fp_nl_params p,
                                                     fp_nl_tables t) {
  __shared__ cpx<double> lds[LdsSize<NT>::value];
  __shared__ double sph[NT];
  constexpr int E = NT / TPB;
  const size_t row = blockIdx.x;
  cpx<double> us[E], k[E];
  double al[E];
  load_row<NT, TPB>(u, row, us);
  split_row<NT, TPB, RS>(us, rho + row * NT, RS ? rhos + row * NT : nullptr, k, al, p, t, lds,
                         sph);
  store_row<NT, TPB>(rhs, row, k);
  if (alpha) {
#pragma unroll
    for (int s = 0; s < E; ++s) alpha[row * NT + threadIdx.x + s * TPB] = al[s];
  }
}

// alpha half-step: u *= exp(-0.5 dz alpha(u, rho))
template <int E, int TPB>
FP_DEV void alpha_half(cpx<double> (&u)[E], const double* __restrict__ rho_row,
                       const fp_nl_params& p, const fp_nl_tables& t) {
  // photo_alpha for the E samples, with its uniform flags tested once outside the
  // per-sample loops (per-sample branches serialised the samples' table loads)
  double a[E];
  if (p.en_pi_loss) {
#pragma unroll
    for (int s = 0; s < E; ++s) {
      const double a2 = u[s].x * u[s].x + u[s].y * u[s].y;
      const double W = keldysh_W(a2, p, t);
      double depl = 1.0 - rho_row[threadIdx.x + s * TPB] / p.rho_max;
      depl = fmin(fmax(depl, 0.0), 1.0);
      a[s] = W * p.photo_pref / (a2 + 1e-30) * depl;
    }
  } else {
#pragma unroll
    for (int s = 0; s < E; ++s) a[s] = 0.0;
  }
  if (p.en_plasma_absorb) {
#pragma unroll
    for (int s = 0; s < E; ++s) a[s] = a[s] + p.plasma_pref * rho_row[threadIdx.x + s * TPB];
  }
#pragma unroll
  for (int s = 0; s < E; ++s) {
    double as = nan0(a[s]);
    as = as > 0.0 ? as : 0.0;
    u[s] = exp(-0.5 * p.dz * as) * u[s];
  }
}

// Nonlinear part of Integrator.step fused over one row. The RK4 base field is
// parked in LDS (stride layout) to keep the register state of a stage small.
template <int NT, int TPB, bool RS>
__global__ __launch_bounds__(TPB) void nl_step_kernel(const double* in, double* out,
                                                       const double* __restrict__ rho,
                                                       const double* __restrict__ rhos,
                                                       int in_spectral, int out_spectral,
                                                       fp_nl_params p, fp_nl_tables t) {
  __shared__ cpx<double> lds[LdsSize<NT>::value];
  __shared__ cpx<double> ub[NT];
  __shared__ double sph[NT];   // 160 KB in all at NT = 4096: one block per CU, as before
  constexpr int E = NT / TPB;
  const int tid = threadIdx.x;
  const size_t row = blockIdx.x;
  const cpx<double>* tw = reinterpret_cast<const cpx<double>*>(t.tw);
  cpx<double> us[E];
  load_row<NT, TPB>(in, row, us);
  if (in_spectral) {
    row_fft<double, NT, TPB, true>(us, lds, tw);
    const double invN = 1.0 / (double)NT;
#pragma unroll
    for (int s = 0; s < E; ++s) us[s] = invN * us[s];
  }
  const double* rr = rho + row * NT;
  const double* rs = RS ? rhos + row * NT : nullptr;
  alpha_half<E, TPB>(us, rr, p, t);
#pragma unroll
  for (int s = 0; s < E; ++s) ub[tid + s * TPB] = us[s];
  const double dz = p.dz;
  cpx<double> acc[E], k[E];
  // classical RK4 with rho frozen: k1 = f(u), k2 = f(u + dz/2 k1), k3 = f(u + dz/2 k2),
  // k4 = f(u + dz k3), u += dz/6 (k1 + 2 k2 + 2 k3 + k4). One split_row body in the
  // instruction stream: the stage loop is not unrolled.
#pragma unroll 1
  for (int stg = 0; stg < 4; ++stg) {
    split_row<NT, TPB, RS>(us, rr, rs, k, nullptr, p, t, lds, sph);
    const double w = (stg == 0 || stg == 3) ? 1.0 : 2.0;
    const double cdz = (stg < 2) ? 0.5 * dz : dz;
#pragma unroll
    for (int s = 0; s < E; ++s) {
      acc[s] = (stg == 0) ? k[s] : acc[s] + w * k[s];
      us[s] = (stg < 3) ? ub[tid + s * TPB] + cdz * k[s] : ub[tid + s * TPB] + (dz / 6.0) * acc[s];
    }
  }
  alpha_half<E, TPB>(us, rr, p, t);
  if (out_spectral) row_fft<double, NT, TPB, false>(us, lds, tw);
  store_row<NT, TPB>(out, row, us);
}

// Record reductions (integrator.py:182-212, operators.py:148-163), one row per block.
template <int TPB>
__global__ __launch_bounds__(TPB) void record_kernel(const double* __restrict__ u,
                                                      const double* __restrict__ rho,
                                                      const double* __restrict__ rhos,
                                                      double* __restrict__ rows_out, fp_nl_params p,
                                                      fp_nl_tables t, double dt, int nbins,
                                                      const int* __restrict__ bin_lo,
                                                      const int* __restrict__ bin_hi,
                                                      double* __restrict__ bins) {
  constexpr int NV = 7;
  __shared__ double red[NV][TPB / 64];
  __shared__ double bred[TPB / 64];
  const int Nt = p.Nt, Nr = p.Nr;
  const size_t row = blockIdx.x;
  const cpx<double>* ur = reinterpret_cast<const cpx<double>*>(u) + row * Nt;
  double v[NV] = {0.0, 0.0, -INFINITY, -INFINITY, 0.0, 0.0, 0.0};
  v[1] = -INFINITY;
  for (int i = threadIdx.x; i < Nt; i += TPB) {
    const cpx<double> z = ur[i];
    const double a2 = z.x * z.x + z.y * z.y;
    const double rr = rho[row * Nt + i];
    const double rs = rhos ? rhos[row * Nt + i] : 0.0;
    const double I = a2 * p.invE2;
    double ph = 0.0;
    if (p.en_pi_loss) {
      const double W = keldysh_W(a2, p, t);
      double depl = fmin(fmax(1.0 - rr / p.rho_max, 0.0), 1.0);
      ph = W * p.photo_pref / (a2 + 1e-30) * depl;
    }
    const double pl = p.en_plasma_absorb ? p.plasma_pref * rr : 0.0;
    v[0] += a2;
    v[1] = fmax(v[1], rr);
    v[2] = fmax(v[2], rs);
    v[3] = fmax(v[3], I);
    v[4] += 2.0 * (ph + pl) * I * dt * 100.0;
    v[5] += 2.0 * ph * I;
    v[6] += 2.0 * pl * I;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double x = v[q];
    for (int d = 32; d >= 1; d >>= 1) {
      const double o = __shfl_xor(x, d, 64);
      x = (q >= 1 && q <= 3) ? fmax(x, o) : x + o;
    }
    if (lane == 0) red[q][wave] = x;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    const int q = threadIdx.x;
    double x = red[q][0];
    for (int w = 1; w < TPB / 64; ++w) x = (q >= 1 && q <= 3) ? fmax(x, red[q][w]) : x + red[q][w];
    rows_out[(size_t)q * Nr + row] = x;
  }
  for (int b = 0; b < nbins; ++b) {
    double x = 0.0;
    for (int i = threadIdx.x; i < Nt; i += TPB) {
      if (i < bin_lo[b] || i >= bin_hi[b]) continue;
      const cpx<double> z = ur[i];
      const double a2 = z.x * z.x + z.y * z.y;
      const double rr = rho[row * Nt + i];
      const double I = a2 * p.invE2;
      double ph = 0.0;
      if (p.en_pi_loss) {
        const double W = keldysh_W(a2, p, t);
        double depl = fmin(fmax(1.0 - rr / p.rho_max, 0.0), 1.0);
        ph = W * p.photo_pref / (a2 + 1e-30) * depl;
      }
      const double pl = p.en_plasma_absorb ? p.plasma_pref * rr : 0.0;
      x += 2.0 * (ph + pl) * I * dt * 100.0;
    }
    for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
    __syncthreads();
    if (lane == 0) bred[wave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
      double s = 0.0;
      for (int w = 0; w < TPB / 64; ++w) s += bred[w];
      bins[(size_t)b * Nr + row] = s;
    }
  }
}

}  // namespace fpnl
