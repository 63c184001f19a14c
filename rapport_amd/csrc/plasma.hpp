// Carrier rate equations (free electrons n_e, self-trapped excitons n_s) along
// t for one radial row, as a parallel multiple-shooting Newton scan.
//
// The recurrence is the exact-exponential update of the reference rate kernel
// (D3/sim/kernels.py:48-105):
//   depl = clip(1 - n_e/n_a, 0, 1)
//   S_e = W*depl + [STE] (W_s + b_s I n_e) n_s/n_a,  L_e = b_g I depl - 1/tau_r
//   S_s = n_e/tau_r,  L_s = -(W_s + b_s I n_e)/n_a - 1/tau_ste
//   n <- exp(L dt) n + S dt phi1(L dt)  (clamped to >= 0, rescaled if n_e+n_s > n_a)
// with I, W, W_s the trapezoid averages of steps k -> k+1.
//
// Each thread owns a chunk of C = NT/TPB consecutive steps. Given a guess of its
// chunk start value it runs the exact sequential recurrence over the chunk (the
// same arithmetic as the reference) together with its tangent (d end / d start,
// scalar without STE, 2x2 with STE). The chunk maps are then linearised around
// the current start values and composed with a workgroup-wide associative scan
// over affine maps x -> A x + b, which yields new chunk start values. The first
// chunk starts from the exact value 0, so iteration i makes chunk i exact and
// the loop ends after at most TPB+1 iterations; the maps are close to affine
// (depletion n/n_a << 1), so in practice it converges in a handful. The first
// guesses are the row's previous densities (the previous z-step's, when the
// row buffer holds them), which makes the first iteration nearly exact. The
// converged starts agree to 1e-14 relative whatever the guess, so results depend
// on the previous buffer contents only at that level. Each iteration writes its
// values to the LDS rows; the last iteration's (from starts within 1e-14 of the
// converged ones) are the result.
#pragma once
#include "common.hpp"
#include "../../include/rapport_amd.h"

namespace fpplasma {

// interpolate_pi_rate (kernels.py:14-26): log10 LUT, linear interpolation.
FP_DEV double lut_rate(double I, const double* __restrict__ tab, const fp_plasma_params& p) {
  if (!fp_isfinite(I) || I < 0.0) I = 0.0;
  if (I < p.Imin) I = p.Imin;
  if (I > p.Imax) I = p.Imax;
  const double fidx = (log10(I) - p.log_min_I) * p.inv_log_step;
  int idx = (int)fidx;
  if (idx < 0) idx = 0;
  if (idx > p.table_size - 2) idx = p.table_size - 2;
  const double r = tab[idx] + (fidx - (double)idx) * (tab[idx + 1] - tab[idx]);
  return (!fp_isfinite(r) || r < 0.0) ? 0.0 : r;
}

struct Step {  // one exact-exponential update with its derivative pieces
  double xnew, E, phi1, dphi1;
};

FP_DEV Step exp_step(double x, double S, double L, double dt) {
  Step s;
  const double Ldt = L * dt;
  s.E = exp(Ldt);
  if (fabs(Ldt) > 1e-6) {
    // one division for both quotients (each then within an ulp of the quotient)
    const double r = 1.0 / Ldt;
    s.phi1 = (s.E - 1.0) * r;
    s.dphi1 = (s.E - s.phi1) * r;
  } else {
    s.phi1 = 1.0 + Ldt * (0.5 + Ldt * (1.0 / 6.0 + Ldt / 24.0));
    s.dphi1 = 0.5 + Ldt * (1.0 / 3.0 + Ldt * 0.125);
  }
  s.xnew = s.E * x + S * dt * s.phi1;
  return s;
}

// State and tangent. Without STE only e/J00 are used.
struct St {
  double e, s;
};
struct Map {  // x -> A x + b
  double a00, a01, a10, a11, b0, b1;
};

template <bool STE>
FP_DEV Map compose(const Map& later, const Map& earlier) {
  Map m;
  if (STE) {
    m.a00 = later.a00 * earlier.a00 + later.a01 * earlier.a10;
    m.a01 = later.a00 * earlier.a01 + later.a01 * earlier.a11;
    m.a10 = later.a10 * earlier.a00 + later.a11 * earlier.a10;
    m.a11 = later.a10 * earlier.a01 + later.a11 * earlier.a11;
    m.b0 = later.a00 * earlier.b0 + later.a01 * earlier.b1 + later.b0;
    m.b1 = later.a10 * earlier.b0 + later.a11 * earlier.b1 + later.b1;
  } else {
    m.a00 = later.a00 * earlier.a00;
    m.b0 = later.a00 * earlier.b0 + later.b0;
    m.a01 = m.a10 = m.a11 = m.b1 = 0.0;
  }
  return m;
}

// One step of the recurrence from (e, s); if TAN, also multiplies the running
// tangent J (row-major 2x2) by the step Jacobian.
// inv_na = 1 / n_a: the ratios n / n_a are products with it (within an ulp of the quotient).
template <bool STE, bool TAN>
FP_DEV St rate_step(St x, double Ia, double Wa, double Wsa, const fp_plasma_params& p,
                    double inv_na, Map& J) {
  const double na = p.na, dt = p.dt;
  const double draw = 1.0 - x.e * inv_na;
  double depl = draw;
  if (depl < 0.0) depl = 0.0;
  if (depl > 1.0) depl = 1.0;
  const double dd = (draw > 0.0 && draw < 1.0) ? -inv_na : 0.0;
  const double wst = STE ? (Wsa + p.beta_s * Ia * x.e) : 0.0;
  const double Se = Wa * depl + (STE ? wst * (x.s * inv_na) : 0.0);
  const double Le = p.beta_g * Ia * depl - p.inv_tau_r;
  const Step se = exp_step(x.e, Se, Le, dt);
  St y;
  y.e = (!fp_isfinite(se.xnew) || se.xnew < 0.0) ? 0.0 : se.xnew;
  const bool ce = (y.e != se.xnew) || !fp_isfinite(se.xnew);
  Step ss;
  double Ss = 0.0, Ls = 0.0;
  bool cs = false;
  if (STE) {
    Ss = p.inv_tau_r * x.e;
    Ls = -wst * inv_na - p.inv_tau_ste;
    ss = exp_step(x.s, Ss, Ls, dt);
    y.s = (!fp_isfinite(ss.xnew) || ss.xnew < 0.0) ? 0.0 : ss.xnew;
    cs = (y.s != ss.xnew) || !fp_isfinite(ss.xnew);
  } else {
    y.s = 0.0;
  }
  // local Jacobian
  double j00 = 0, j01 = 0, j10 = 0, j11 = 0;
  if (TAN) {
    const double dLe_de = p.beta_g * Ia * dd;
    const double dSe_de = Wa * dd + (STE ? p.beta_s * Ia * (x.s * inv_na) : 0.0);
    j00 = se.E + x.e * se.E * dt * dLe_de + dt * (dSe_de * se.phi1 + Se * se.dphi1 * dt * dLe_de);
    if (STE) {
      j01 = dt * (wst * inv_na) * se.phi1;
      const double dLs_de = -p.beta_s * Ia * inv_na;
      j10 = x.s * ss.E * dt * dLs_de + dt * (p.inv_tau_r * ss.phi1 + Ss * ss.dphi1 * dt * dLs_de);
      j11 = ss.E;
    }
    if (ce) { j00 = 0.0; j01 = 0.0; }
    if (cs) { j10 = 0.0; j11 = 0.0; }
  }
  const double tot = y.e + y.s;
  if (tot > na) {
    const double scale = na / tot;
    if (TAN) {
      const double f = scale / tot;
      const double r00 = f * y.s, r01 = -f * y.e;
      const double r10 = -f * y.s, r11 = f * y.e;
      const double n00 = r00 * j00 + r01 * j10, n01 = r00 * j01 + r01 * j11;
      const double n10 = r10 * j00 + r11 * j10, n11 = r10 * j01 + r11 * j11;
      j00 = n00; j01 = n01; j10 = n10; j11 = n11;
    }
    y.e *= scale;
    y.s *= scale;
  }
  if (TAN) {
    Map L;
    L.a00 = j00; L.a01 = j01; L.a10 = j10; L.a11 = j11; L.b0 = L.b1 = 0.0;
    J = compose<STE>(L, J);
  }
  return y;
}

// LDS needed by plasma_row (bytes): one NT-double row (two with STE: n_e and n_s) + scan
// scratch. Without STE a 256-thread row then takes 37 KB, so four rows fit in one CU's LDS.
template <int NT, int TPB, bool STE>
struct PlasmaLds {
  static constexpr int rows = STE ? 2 : 1;
  static constexpr int bytes = (rows * NT + 2 * (TPB + 1) + 6 * (TPB / 64)) * 8;
};

// Solves one row. On entry sI[0..NT) holds I_k = |E_k|^2 * field_to_I in LDS
// (all threads' writes visible). Writes n_e, n_s of the row to global memory
// (ne_row, ns_row; ns_row may be null when !STE). `smem` is the PlasmaLds area
// whose first NT doubles are sI. Returns the number of Newton iterations.
template <int NT, int TPB, bool STE>
FP_DEV int plasma_row(double* smem, const fp_plasma_params& p, const double* __restrict__ Wg,
                      const double* __restrict__ Ws, double* __restrict__ ne_row,
                      double* __restrict__ ns_row, int* status) {
  constexpr int C = NT / TPB;
  constexpr int NW = TPB / 64;
  double* sI = smem;
  double* sS = STE ? smem + NT : nullptr;                 // second row buffer (n_s output)
  double* sb0 = smem + PlasmaLds<NT, TPB, STE>::rows * NT;  // chunk start offsets (TPB+1)
  double* sb1 = sb0 + (TPB + 1);
  double* swt = sb1 + (TPB + 1);          // wave totals, 6 per wave
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int k0 = tid * C;
  const int nsteps = (tid == TPB - 1) ? C - 1 : C;   // steps k0 .. k0+nsteps-1
  const double inv_na = 1.0 / p.na;

  double Ia[C], Wa[C], Wsa[C];
  {
    double Ic = sI[k0];
    double Wc = lut_rate(Ic, Wg, p);
    double Wsc = STE ? lut_rate(Ic, Ws, p) : 0.0;
#pragma unroll
    for (int i = 0; i < C; ++i) {
      const int kn = (k0 + i + 1 < NT) ? k0 + i + 1 : NT - 1;
      const double In = sI[kn];
      const double Wn = lut_rate(In, Wg, p);
      const double Wsn = STE ? lut_rate(In, Ws, p) : 0.0;
      Ia[i] = 0.5 * (Ic + In);
      Wa[i] = 0.5 * (Wc + Wn);
      Wsa[i] = 0.5 * (Wsc + Wsn);
      Ic = In; Wc = Wn; Wsc = Wsn;
    }
  }

  __syncthreads();   // every thread has read its intensities before the rows are overwritten

  // First guess of this chunk's start value: the row's previous content (the
  // densities of the previous z-step, which the field changes little), else 0.
  St start;
  start.e = 0.0;
  start.s = 0.0;
  if (k0 > 0) {
    const double ge = ne_row[k0];
    if (ge >= 0.0 && ge <= p.na) start.e = ge;
    if (STE && ns_row) {
      const double gs = ns_row[k0];
      if (gs >= 0.0 && gs <= p.na) start.s = gs;
    }
  }
  int it = 0;
  const int maxit = TPB + 2;
  for (;;) {
    ++it;
    // integrate the chunk with tangent
    // (the values go to the LDS rows as they are produced: once the starts have
    // converged, the last iteration's values are the row -- no separate final pass)
    Map J;
    J.a00 = 1.0; J.a01 = 0.0; J.a10 = 0.0; J.a11 = 1.0; J.b0 = J.b1 = 0.0;
    St x = start;
#pragma unroll
    for (int i = 0; i < C; ++i) {
      if (i < nsteps) {
        x = rate_step<STE, true>(x, Ia[i], Wa[i], Wsa[i], p, inv_na, J);
        sI[k0 + i + 1] = x.e;
        if (STE) sS[k0 + i + 1] = x.s;
      }
    }
    // affine map of chunk: s_{c+1} = J (s_c - start) + x
    Map m = J;
    m.b0 = x.e - (J.a00 * start.e + (STE ? J.a01 * start.s : 0.0));
    m.b1 = STE ? x.s - (J.a10 * start.e + J.a11 * start.s) : 0.0;
    // inclusive wave scan
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      Map o;
      o.a00 = __shfl_up(m.a00, d, 64);
      o.b0 = __shfl_up(m.b0, d, 64);
      if (STE) {
        o.a01 = __shfl_up(m.a01, d, 64);
        o.a10 = __shfl_up(m.a10, d, 64);
        o.a11 = __shfl_up(m.a11, d, 64);
        o.b1 = __shfl_up(m.b1, d, 64);
      } else {
        o.a01 = o.a10 = o.a11 = o.b1 = 0.0;
      }
      if (lane >= d) m = compose<STE>(m, o);
    }
    if (lane == 63) {
      swt[6 * wave + 0] = m.a00; swt[6 * wave + 1] = m.a01; swt[6 * wave + 2] = m.a10;
      swt[6 * wave + 3] = m.a11; swt[6 * wave + 4] = m.b0; swt[6 * wave + 5] = m.b1;
    }
    __syncthreads();
    if (wave > 0) {
      Map pre;
      pre.a00 = swt[0]; pre.a01 = swt[1]; pre.a10 = swt[2]; pre.a11 = swt[3];
      pre.b0 = swt[4]; pre.b1 = swt[5];
      for (int w = 1; w < wave; ++w) {
        Map t;
        t.a00 = swt[6 * w]; t.a01 = swt[6 * w + 1]; t.a10 = swt[6 * w + 2];
        t.a11 = swt[6 * w + 3]; t.b0 = swt[6 * w + 4]; t.b1 = swt[6 * w + 5];
        pre = compose<STE>(t, pre);
      }
      m = compose<STE>(m, pre);
    }
    // start of chunk c+1 = F_c(0) = b
    sb0[tid + 1] = m.b0;
    if (STE) sb1[tid + 1] = m.b1;
    if (tid == 0) { sb0[0] = 0.0; sb1[0] = 0.0; }
    __syncthreads();
    St ns_;
    ns_.e = sb0[tid];
    ns_.s = STE ? sb1[tid] : 0.0;
    const double de = fabs(ns_.e - start.e), ds = fabs(ns_.s - start.s);
    const bool ok = (de <= 1e-14 * fabs(ns_.e) + 1e-300) &&
                    (!STE || ds <= 1e-14 * fabs(ns_.s) + 1e-300);
    start = ns_;
    const int all_ok = __syncthreads_and(ok ? 1 : 0);
    if (all_ok || it >= maxit) {
      if (!all_ok && tid == 0 && status) atomicAdd(status, 1);
      break;
    }
  }

  if (k0 == 0) { sI[0] = 0.0; if (STE) sS[0] = 0.0; }
  __syncthreads();
  for (int k = tid; k < NT; k += TPB) {
    ne_row[k] = sI[k];
    if (ns_row) ns_row[k] = STE ? sS[k] : 0.0;
  }
  return it;
}

}  // namespace fpplasma
