// 1-D SPM + optical-shock model (propagate_shock, D3/sim/self_steepening.py:80-105;
// right-hand side shock_rhs :65-77, spectral derivative _dt_spectral :60-62):
//   dE/dz = i spm |E|^2 E - shock ifft(i w fft(|E|^2 E)),   w = 2 pi fftfreq(N, dt),
// classical RK4 in z, FP64.
//
// One workgroup owns one pulse for the whole z march: the field, the RK4 sum and the
// stage field stay in registers in the FFT stride layout (element tid + s TPB), each
// right-hand side is one forward and one inverse LDS Stockham transform, and nothing
// but the saved planes goes to HBM. For N = 8192 (1024 threads, 128 registers each) the
// step's base field, the RK4 sum and the stage field live in three global scratch rows
// instead (each thread reads and writes only its own elements, so no barrier is needed;
// the rows stay L2 resident), and only the transform buffer is held in registers.
#pragma once
#include "common.hpp"
#include "fft.hpp"
#include "../../include/rapport_amd.h"

namespace fpshock {

typedef cpx<double> c128;

// angular frequency of FFT bin k: 2 pi * (kk / (N dt)) with kk = k or k - N (numpy's
// fftfreq: the integer index times 1 / (N d), then the 2 pi factor)
FP_DEV double omega_of(int k, int N, double val) {
  const int kk = k < (N + 1) / 2 ? k : k - N;
  return 6.283185307179586 * ((double)kk * val);
}

// k = i spm IE - shock ifft(i w fft(IE)), IE = |us|^2 us, for one row; the stage field
// us comes from registers (us[s]) or, with UG, from the thread's own elements of the
// global row wu (then only k is live across the transforms).
template <int N, int TPB, bool UG>
FP_DEV void shock_rhs(const c128 (&us)[UG ? 1 : N / TPB], const c128* wu, c128 (&k)[N / TPB],
                      c128* lds, const c128* __restrict__ tw, const fp_shock_row& P, double val,
                      int tid) {
  constexpr int E = N / TPB;
  auto uget = [&](int s) -> c128 { return UG ? wu[tid + s * TPB] : us[UG ? 0 : s]; };
#pragma unroll
  for (int s = 0; s < E; ++s) {
    const c128 u = uget(s);
    const double I = u.x * u.x + u.y * u.y;
    k[s] = I * u;
  }
  if (P.shock != 0.0) {
    fpfft::row_fft<double, N, TPB, false>(k, lds, tw);
#pragma unroll
    for (int s = 0; s < E; ++s) {
      const double w = omega_of(tid + s * TPB, N, val);
      k[s] = mk<double>(-w * k[s].y, w * k[s].x);              // (i w) F
    }
    fpfft::row_fft<double, N, TPB, true>(k, lds, tw);
  }
  const double invN = 1.0 / (double)N;
#pragma unroll
  for (int s = 0; s < E; ++s) {
    const c128 u = uget(s);
    const double I = u.x * u.x + u.y * u.y;
    const c128 IE = I * u;
    c128 r = mk<double>(-P.spm * IE.y, P.spm * IE.x);         // i spm IE
    if (P.shock != 0.0) {
      const c128 d = invN * k[s];                              // ifft scaling
      r = r - P.shock * d;
    }
    k[s] = r;
  }
}

// UG: the step's base field, the RK4 sum and the stage field in the global scratch rows
// `work` (3 N complex per row) instead of registers.
template <int N, int TPB, bool UG>
__global__ __launch_bounds__(TPB) void shock_kernel(const double* __restrict__ E0,
                                                    double* __restrict__ out,
                                                    double* __restrict__ work,
                                                    const fp_shock_row* __restrict__ rows, int nz,
                                                    int stride, int n_saves,
                                                    const double* __restrict__ twd) {
  __shared__ c128 lds[fpfft::LdsSize<N>::value];
  constexpr int E = N / TPB;
  constexpr int RE = UG ? 1 : E;      // register copies of base, sum and stage field
  const int tid = threadIdx.x;
  const size_t row = blockIdx.x;
  const fp_shock_row P = rows[row];
  const double val = 1.0 / ((double)N * P.dt);
  const c128* tw = reinterpret_cast<const c128*>(twd);
  const c128* e0 = reinterpret_cast<const c128*>(E0) + row * N;
  c128* o = reinterpret_cast<c128*>(out) + row * (size_t)n_saves * N;
  c128* wb = UG ? reinterpret_cast<c128*>(work) + row * 3 * N : nullptr;   // base row
  c128* wa = UG ? wb + N : nullptr;                                        // RK4 sum row
  c128* wu = UG ? wb + 2 * N : nullptr;                                    // stage field row

  c128 ureg[RE], areg[RE], us[RE];
  c128 k[E];
#pragma unroll
  for (int s = 0; s < E; ++s) {
    const int e = tid + s * TPB;
    const c128 v = e0[e];
    o[e] = v;
    if (UG) {
      wb[e] = v;
      wu[e] = v;
    } else {
      ureg[UG ? 0 : s] = v;
      us[UG ? 0 : s] = v;
    }
  }

  const double h = 0.5 * P.dz, d6 = P.dz / 6.0;
  int save = 1;
  for (int i = 0; i < nz; ++i) {
#pragma unroll 1
    for (int stg = 0; stg < 4; ++stg) {
      // opaque copy of the thread index: keeps the per-element scratch addresses and
      // frequencies from being hoisted out of the z loop into registers
      int ot = tid;
      asm volatile("" : "+v"(ot));
      shock_rhs<N, TPB, UG>(us, wu, k, lds, tw, P, val, ot);
#pragma unroll
      for (int s = 0; s < E; ++s) {
        const int e = ot + s * TPB;
        c128 a = stg == 0 ? k[s] : (UG ? wa[e] : areg[UG ? 0 : s]);
        if (stg == 1 || stg == 2) a = a + 2.0 * k[s];
        else if (stg == 3) a = a + k[s];
        const c128 u = UG ? wb[e] : ureg[UG ? 0 : s];
        const c128 nu = stg < 2 ? u + h * k[s] : (stg == 2 ? u + P.dz * k[s] : u + d6 * a);
        if (UG) {
          wa[e] = a;
          wu[e] = nu;
          if (stg == 3) wb[e] = nu;
        } else {
          areg[UG ? 0 : s] = a;
          us[UG ? 0 : s] = nu;
          if (stg == 3) ureg[UG ? 0 : s] = nu;
        }
      }
    }
    // the stage field now holds the new field
    if ((i + 1) % stride == 0 && save < n_saves) {
      c128* dst = o + (size_t)save * N;
      int ot = tid;
      asm volatile("" : "+v"(ot));
#pragma unroll
      for (int s = 0; s < E; ++s) {
        const int e = ot + s * TPB;
        dst[e] = UG ? wu[e] : us[UG ? 0 : s];
      }
      ++save;
    }
  }
}

}  // namespace fpshock
