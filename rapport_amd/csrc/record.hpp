// One recording of the propagate() loop (_record, D3/sim/integrator.py:179-224, with
// loss_rates, D3/sim/operators.py:148-163) in two launches, straight from the spectral
// state into the float32 histories.
//
// record_history_kernel: one workgroup per radial row, eight samples per thread. It reads
// the row of the spectral state (complex128, or the mixed path's complex64), forms the
// time-domain row with an inverse LDS transform and the 1/NT scale of fft_rows_kernel (the
// complex64 row is upcast after the scale), and reduces in FP64: sum_t |u|^2, max_t rho,
// max_t rho_s, max_t I, sum_t dep, sum_t 2 photo I, sum_t 2 plasma I, with
// I = |u|^2 invE2 and dep = 2 (photo + plasma) I dt 100. It writes
// this record's history entries as float32: fluence (sum |u|^2 dt invE2), rho, rho_s, peak
// intensity, absorbed, the absorbed time bins, the density column at the snapshot time,
// and, for row 0, the on-axis rows and, for rows on the radial stride, the subsampled
// (r, t) cube. The three quantities taken over rows (max I, and the two loss integrals)
// go to a 3 x Nr FP64 scratch.
//
// record_finish_kernel: one workgroup reduces that scratch in a fixed order into the peak
// intensity and the two dE/dz entries (sum_r loss_r dt w_r 1e4 1e6, uJ/m) of the record.
// No float atomics: repeated runs give identical histories.
#pragma once
#include "../../include/rapport_amd.h"
#include "common.hpp"
#include "fft.hpp"
#include "nl.hpp"

namespace fprec {

template <typename T, int NT, int TPB>
__global__ __launch_bounds__(TPB) void record_history_kernel(
    const T* __restrict__ spec, const T* __restrict__ tw, const double* __restrict__ rho,
    const double* __restrict__ rhos, fp_nl_params p, fp_nl_tables t, double dt, int nbins,
    const int* __restrict__ bin_lo, const int* __restrict__ bin_hi, fp_record_dst d) {
  constexpr int E = NT / TPB, NW = TPB / 64;
  __shared__ cpx<T> lds[fpfft::LdsSize<NT>::value];
  __shared__ double red[8][NW];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int Nr = p.Nr;
  const size_t row = blockIdx.x;

  cpx<T> v[E];
  const cpx<T>* sp = reinterpret_cast<const cpx<T>*>(spec) + row * NT;
#pragma unroll
  for (int s = 0; s < E; ++s) v[s] = sp[tid + s * TPB];
  // the densities are loaded before the transform, so their latency overlaps it (loads are
  // not moved across the transform's barriers)
  const double* rr_row = rho + row * NT;
  const double* rs_row = rhos ? rhos + row * NT : nullptr;
  double rrv[E], rsv[E];
#pragma unroll
  for (int s = 0; s < E; ++s) {
    rrv[s] = rr_row[tid + s * TPB];
    rsv[s] = rs_row ? rs_row[tid + s * TPB] : 0.0;
  }
  fpfft::row_fft<T, NT, TPB, true>(v, lds, reinterpret_cast<const cpx<T>*>(tw));
  const T sc = T(1) / (T)NT;

  double vs[7] = {0.0, -INFINITY, -INFINITY, -INFINITY, 0.0, 0.0, 0.0};
  double dep[E];
  const bool cube_row = d.t_stride > 0 && (int)(row % (size_t)d.r_stride) == 0;
  const size_t cube_off = cube_row ? (row / d.r_stride) * (size_t)d.Nt_sub : 0;
#pragma unroll
  for (int s = 0; s < E; ++s) {
    const int i = tid + s * TPB;
    const cpx<T> w = sc * v[s];
    const double x = (double)w.x, y = (double)w.y;
    const double a2 = x * x + y * y;
    const double rr = rrv[s];
    const double rs = rsv[s];
    const double I = a2 * p.invE2;
    double ph = 0.0;
    if (p.en_pi_loss) {
      const double W = fpnl::keldysh_W(a2, p, t);
      const double depl = fmin(fmax(1.0 - rr / p.rho_max, 0.0), 1.0);
      ph = W * p.photo_pref / (a2 + 1e-30) * depl;
    }
    const double pl = p.en_plasma_absorb ? p.plasma_pref * rr : 0.0;
    dep[s] = 2.0 * (ph + pl) * I * dt * 100.0;
    vs[0] += a2;
    vs[1] = fmax(vs[1], rr);
    vs[2] = fmax(vs[2], rs);
    vs[3] = fmax(vs[3], I);
    vs[4] += dep[s];
    vs[5] += 2.0 * ph * I;
    vs[6] += 2.0 * pl * I;
    if (row == 0) {
      d.rho_onaxis[i] = (float)rr;
      d.rho_s_onaxis[i] = (float)rs;
      d.I_onaxis[i] = (float)I;
    }
    if (cube_row && i % d.t_stride == 0) {
      const size_t o = cube_off + i / d.t_stride;
      d.rho_cube[o] = (float)rr;
      d.rho_s_cube[o] = (float)rs;
      d.I_cube[o] = (float)I;
    }
  }
#pragma unroll
  for (int q = 0; q < 7; ++q) {
    double x = vs[q];
    for (int dd = 32; dd >= 1; dd >>= 1) {
      const double o = __shfl_xor(x, dd, 64);
      x = (q >= 1 && q <= 3) ? fmax(x, o) : x + o;
    }
    if (lane == 0) red[q][wave] = x;
  }
  __syncthreads();
  if (tid < 7) {
    const int q = tid;
    double x = red[q][0];
    for (int w = 1; w < NW; ++w) x = (q >= 1 && q <= 3) ? fmax(x, red[q][w]) : x + red[q][w];
    switch (q) {
      case 0: d.fluence[row] = (float)(x * dt * p.invE2); break;
      case 1: d.rho[row] = (float)x; break;
      case 2: d.rho_s[row] = (float)x; break;
      case 3: d.ipeak[row] = (float)x; d.partial[row] = x; break;
      case 4: d.absorbed[row] = (float)x; break;
      case 5: d.partial[Nr + row] = x; break;
      default: d.partial[2 * Nr + row] = x; break;
    }
  }
  if (d.rho_at && tid == 0) d.rho_at[row] = (float)rr_row[d.it_snap];
  for (int b = 0; b < nbins; ++b) {
    const int lo = bin_lo[b], hi = bin_hi[b];
    double x = 0.0;
#pragma unroll
    for (int s = 0; s < E; ++s) {
      const int i = tid + s * TPB;
      if (i >= lo && i < hi) x += dep[s];
    }
    for (int dd = 32; dd >= 1; dd >>= 1) x += __shfl_xor(x, dd, 64);
    __syncthreads();
    if (lane == 0) red[7][wave] = x;
    __syncthreads();
    if (tid == 0) {
      double sum = red[7][0];
      for (int w = 1; w < NW; ++w) sum += red[7][w];
      d.bins[(size_t)b * Nr + row] = (float)sum;
    }
  }
}

constexpr int FIN_TPB = 256;

__global__ __launch_bounds__(FIN_TPB) void record_finish_kernel(int Nr, double dt, fp_record_dst d) {
  constexpr int NW = FIN_TPB / 64;
  __shared__ double red[3][NW];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  double mx = -INFINITY, s1 = 0.0, s2 = 0.0;
  for (int r = tid; r < Nr; r += FIN_TPB) {
    mx = fmax(mx, d.partial[r]);
    s1 += d.partial[Nr + r] * dt * d.r_w[r];
    s2 += d.partial[2 * Nr + r] * dt * d.r_w[r];
  }
  for (int dd = 32; dd >= 1; dd >>= 1) {
    mx = fmax(mx, __shfl_xor(mx, dd, 64));
    s1 += __shfl_xor(s1, dd, 64);
    s2 += __shfl_xor(s2, dd, 64);
  }
  if (lane == 0) {
    red[0][wave] = mx;
    red[1][wave] = s1;
    red[2][wave] = s2;
  }
  __syncthreads();
  if (tid == 0) {
    for (int w = 1; w < NW; ++w) {
      mx = fmax(mx, red[0][w]);
      s1 += red[1][w];
      s2 += red[2][w];
    }
    *d.imax = mx;
    *d.dedz_photo = s1 * 1e4 * 1e6;
    *d.dedz_plasma = s2 * 1e4 * 1e6;
  }
}

}  // namespace fprec
