// 0-D re-integration of the carrier rate equations over given intensity histories
// (integre_0d, D3/sim/figures_article.py:168-194; _phi1 / _step :161-166).
//
// Each row is an independent sequential recursion over its Nt samples; a thread owns
// one row and a block of 64 threads owns 64 rows. Rows are staged through LDS in
// chunks of TCH samples so that every global load and store is a contiguous run along
// t (a thread reading its own row directly would stride by Nt across the wave). Within
// a chunk the rates W(I), W_s(I) of all 64 x TCH samples are evaluated first, in
// parallel (independent PCHIP evaluations that overlap their table loads), and the
// sequential pass then reads only LDS.
//
// Semantics follow the reference step for step: averages of I, W, W_s over the two
// samples of a step; depletion clip(1 - n_e/n_a, 0, 1) from the start of the step;
// n <- e^{L dt} n + S dt phi1(L dt), replaced by 0 when not finite or negative; both
// densities rescaled by n_a / (n_e + n_s) when the sum exceeds n_a. NaN intensities
// propagate into the rates as in NumPy (comparisons, not fmin/fmax, do the clipping).
#pragma once
#include "common.hpp"
#include "../../include/rapport_amd.h"

namespace fprate0d {

constexpr int ROWS = 64;   // rows (threads) per block
constexpr int TCH = 16;    // samples per staged chunk

// phi1(x) = expm1(x) / x, with the reference's series below |x| = 1e-6
FP_DEV double phi1(double x) {
  return fabs(x) > 1e-6 ? expm1(x) / x : 1.0 + x * (0.5 + x * (1.0 / 6.0 + x / 24.0));
}

FP_DEV double exp_step(double x, double S, double L, double dt) {
  const double xn = exp(L * dt) * x + S * dt * phi1(L * dt);
  return (fp_isfinite(xn) && xn >= 0.0) ? xn : 0.0;
}

// max(pchip(log10(clip(I, Imin, Imax))), 0) with scipy PPoly semantics: interval i with
// knots[i] <= x < knots[i+1] (the last interval for x >= the last knot), value
// c3 + c2 s + c1 s^2 + c0 s^3 summed in that order, s = x - knots[i]. The first guess
// assumes near-uniform knots; the two loops make the choice exact for any increasing set.
FP_DEV double pchip_rate(double I, const double* __restrict__ knots, const double* __restrict__ c,
                         int n, double inv_h, double Imin, double Imax) {
  double Ic = I < Imin ? Imin : I;
  Ic = Ic > Imax ? Imax : Ic;
  const double x = log10(Ic);
  if (!(x == x)) return x;   // NaN in, NaN out (NumPy's maximum keeps NaN)
  const int m = n - 1;
  double g = (x - knots[0]) * inv_h;
  g = g >= 0.0 ? g : 0.0;                       // also NaN (degenerate knots) -> 0
  g = g > (double)(m - 1) ? (double)(m - 1) : g;
  int i = (int)g;
  while (i > 0 && knots[i] > x) --i;
  while (i < m - 1 && knots[i + 1] <= x) ++i;
  const double s = x - knots[i];
  double r = c[3 * m + i], z = s;
  r += c[2 * m + i] * z;
  z *= s;
  r += c[m + i] * z;
  z *= s;
  r += c[i] * z;
  return r < 0.0 ? 0.0 : r;
}

struct Smem {
  double I[ROWS][TCH + 1];
  double Wg[ROWS][TCH + 1];
  double Ws[ROWS][TCH + 1];
  double ne[ROWS][TCH + 1];
  double ns[ROWS][TCH + 1];
};

__global__ __launch_bounds__(ROWS) void rate0d_kernel(const double* __restrict__ t,
                                                      const double* __restrict__ I,
                                                      double* __restrict__ ne_out,
                                                      double* __restrict__ ns_out,
                                                      fp_rate0d_params p,
                                                      const double* __restrict__ knots,
                                                      const double* __restrict__ cg,
                                                      const double* __restrict__ cs) {
  __shared__ Smem sm;
  const int lane = threadIdx.x;
  const int row0 = blockIdx.x * ROWS;
  const int Nt = p.Nt;
  const bool live = row0 + lane < p.n_rows;
  const double inv_h = (double)(p.n_knots - 1) / (knots[p.n_knots - 1] - knots[0]);
  const double na = p.na, inv_tr = p.inv_tau_r;

  // carried state: densities and the previous sample's (I, W, W_s)
  double ne = 0.0, ns = 0.0, Ip = 0.0, Wp = 0.0, Wsp = 0.0;
  for (int k0 = 0; k0 < Nt; k0 += TCH) {
    const int cnt = Nt - k0 < TCH ? Nt - k0 : TCH;
    // stage the chunk and evaluate its rates (64 x TCH independent samples)
    for (int e = lane; e < ROWS * TCH; e += ROWS) {
      const int r = e / TCH, j = e % TCH;
      if (row0 + r < p.n_rows && j < cnt) {
        double v = I[(size_t)(row0 + r) * Nt + k0 + j];
        v = v < 0.0 ? 0.0 : v;                       // np.clip(I, 0, None)
        sm.I[r][j] = v;
        sm.Wg[r][j] = pchip_rate(v, knots, cg, p.n_knots, inv_h, p.Imin, p.Imax);
        sm.Ws[r][j] = p.ste ? pchip_rate(v, knots, cs, p.n_knots, inv_h, p.Imin, p.Imax) : 0.0;
      }
    }
    __syncthreads();
    if (live) {
      for (int j = 0; j < cnt; ++j) {
        const int k = k0 + j;
        const double Ik = sm.I[lane][j], Wk = sm.Wg[lane][j], Wsk = sm.Ws[lane][j];
        if (k > 0) {
          const double dt = t[k] - t[k - 1];
          const double Iavg = 0.5 * (Ip + Ik);
          const double Wavg = 0.5 * (Wp + Wk);
          const double Wsav = 0.5 * (Wsp + Wsk);
          double depl = 1.0 - ne / na;
          depl = depl < 0.0 ? 0.0 : (depl > 1.0 ? 1.0 : depl);
          const double ion_s = Wsav + p.beta_s * Iavg * ne;   // STE re-ionisation rate
          const double Se = Wavg * depl + (p.ste ? ion_s * (ns / na) : 0.0);
          const double Le = (p.avalanche ? p.beta_g * Iavg * depl : 0.0) - (p.trapping ? inv_tr : 0.0);
          const double Ss = (p.ste && p.trapping) ? inv_tr * ne : 0.0;
          const double Ls = p.ste ? -ion_s / na : 0.0;
          double ne_n = exp_step(ne, Se, Le, dt);
          double ns_n = p.ste ? exp_step(ns, Ss, Ls, dt) : 0.0;
          const double tot = ne_n + ns_n;
          if (tot > na) {
            const double f = na / tot;
            ne_n *= f;
            ns_n *= f;
          }
          ne = ne_n;
          ns = ns_n;
        }
        sm.ne[lane][j] = ne;
        sm.ns[lane][j] = ns;
        Ip = Ik;
        Wp = Wk;
        Wsp = Wsk;
      }
    }
    __syncthreads();
    for (int e = lane; e < ROWS * TCH; e += ROWS) {
      const int r = e / TCH, j = e % TCH;
      if (row0 + r < p.n_rows && j < cnt) {
        const size_t o = (size_t)(row0 + r) * Nt + k0 + j;
        ne_out[o] = sm.ne[r][j];
        ns_out[o] = sm.ns[r][j];
      }
    }
    __syncthreads();
  }
}

}  // namespace fprate0d
