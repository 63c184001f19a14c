// Raman convolution of split() (D3/sim/operators.py:109-112):
//   K_R = Re(IFFT(FFT(I) R_f)),  R_f = dt FFT(ifftshift(R_t)),
//   R_t(t) = Omega_r2 tau_s exp(-t/tau_d) sin(t/tau_s) for t > 0, 0 otherwise
// (D3/sim/grids.py:113-116). This is the circular convolution
//   K_R[n] = sum_{k=1}^{N/2-1} h[k] I[(n - k) mod N],  h[k] = C Im(w^k),
//   w = exp(dt (-1/tau_d + i/tau_s)),  C = dt Omega_r2 tau_s,
// which equals, exactly, K_R[n] = Im(P[n] - w^{N/2} (I'[n - N/2] + P[n - N/2])) with
// I' = C I and P the periodic solution of the first-order recursion
//   P[n] = w (P[n-1] + I'[n-1])   (P[0] = P[N]):
// P sums all lags k >= 1 of the periodic signal, and the w^{N/2} term removes the
// lags k >= N/2 that the sampled kernel does not have.
//
// The recursion runs as a workgroup scan: each thread folds a contiguous chunk of E
// samples into one affine map P -> w^E P + B (the multiplier is the same for every
// chunk), a weighted inclusive scan combines the chunks (wave shuffles, then wave
// totals through LDS, carried across waves by the same recursion), the periodic start is
// P[0] = X_total / (1 - w^N), and each
// thread regenerates its chunk. Input and output use the FFT stride layout
// (v[s] = x[tid + s TPB]); sample n - N/2 is then register (s + E/2) mod E of the
// same thread. Replaces a forward and an inverse row FFT.
//
// wt: N + 2 complex entries, wt[m] = w^m (m = 0..N), wt[N + 1] = (C, 0).
#pragma once
#include "common.hpp"
#include "fft.hpp"

namespace fpraman {

template <typename T>
FP_DEV cpx<T> shfl_up_c(cpx<T> v, int d) {
  return mk<T>(__shfl_up(v.x, d, 64), __shfl_up(v.y, d, 64));
}

template <typename T>
FP_DEV cpx<T> cdiv(cpx<T> a, cpx<T> b) {
  const T den = b.x * b.x + b.y * b.y;
  return mk<T>((a.x * b.x + a.y * b.y) / den, (a.y * b.x - a.x * b.y) / den);
}

// `lds` must hold fpfft::LdsSize<NT>::value complex entries and be free on entry; it
// is free again on return (the routine ends with a barrier).
template <typename T, int NT, int TPB>
FP_DEV void raman_row(const cpx<T> (&us)[NT / TPB], T (&KR)[NT / TPB], cpx<T>* lds,
                      const cpx<T>* __restrict__ wt) {
  constexpr int E = NT / TPB;
  static_assert(E % 2 == 0, "the N/2 shift must stay inside a thread");
  static_assert(TPB % 64 == 0, "whole waves");
  static_assert((E & (E - 1)) == 0, "samples per thread: a power of two");
  int tid = threadIdx.x & (TPB - 1);
  asm volatile("" : "+v"(tid));
  const int lane = tid & 63, wave = tid >> 6;
  const T C = wt[NT + 1].x;
  const cpx<T> w = wt[1];
  T* lf = reinterpret_cast<T*>(lds);          // real staging, index n + n/8 (< 2 NT reals)

  // I' = C |u|^2 in the stride layout, staged for the chunk layout
  T Is[E];
#pragma unroll
  for (int s = 0; s < E; ++s) {
    const int n = tid + s * TPB;
    Is[s] = C * (us[s].x * us[s].x + us[s].y * us[s].y);
    lf[n + (n >> 3)] = Is[s];
  }
  __syncthreads();
  T Ic[E];
  cpx<T> B = mk<T>(T(0), T(0));
  // with E a multiple of 8, (tid E + j) >> 3 = (tid E) >> 3 + (j >> 3): one base per thread
  const int cb = (E % 8 == 0) ? tid * E + ((tid * E) >> 3) : 0;
#pragma unroll
  for (int j = 0; j < E; ++j) {
    const int n = tid * E + j;
    Ic[j] = (E % 8 == 0) ? lf[cb + j + (j >> 3)] : lf[n + (n >> 3)];
    B = w * mk<T>(B.x + Ic[j], B.y);          // B = sum_j w^(E-j) I'_j
  }
  // weighted inclusive scan X_t = sum_{u <= t} w^(E (t - u)) B_u
  cpx<T> X = B;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const cpx<T> o = shfl_up_c(X, d);
    if (lane >= d) X = X + wt[E * d] * o;
  }
  __syncthreads();                            // chunk reads of lf done
  if (lane == 63) lds[wave] = X;
  __syncthreads();
  // carry from the earlier waves: G = scan value at the last thread of wave - 1, built
  // from the wave totals T_v by G_v = T_v + w^(64 E) G_{v-1} (broadcast LDS reads), then
  // X += w^(E (lane + 1)) G -- one per-lane table read instead of one per earlier wave
  if (wave > 0) {
    const cpx<T> w64 = wt[64 * E];
    cpx<T> G = lds[0];
    for (int v = 1; v < wave; ++v) G = lds[v] + w64 * G;
    X = X + wt[E * (lane + 1)] * G;
  }
  __syncthreads();                            // wave totals read
  lds[tid] = X;
  __syncthreads();
  const cpx<T> XT = lds[TPB - 1];
  cpx<T> P = tid > 0 ? lds[tid - 1] : mk<T>(T(0), T(0));
  const cpx<T> wN = wt[NT];
  const cpx<T> P0 = cdiv(XT, mk<T>(T(1) - wN.x, -wN.y));
  P = P + wt[E * tid] * P0;                   // + w^(E t) P[0]
  __syncthreads();                            // prefix values read
  // tid E and j < E occupy disjoint bits (E a power of two), so their sum is their XOR,
  // and lds_pad is linear over XOR
  const int pb = fpfft::lds_pad(tid * E);
#pragma unroll
  for (int j = 0; j < E; ++j) {
    lds[pb ^ fpfft::lds_pad(j)] = P;
    P = w * mk<T>(P.x + Ic[j], P.y);
  }
  __syncthreads();
  cpx<T> Ps[E];
  if constexpr (TPB % 256 == 0) {
    const int b = fpfft::lds_pad(tid);          // as in the transforms' stride-layout reads
#pragma unroll
    for (int s = 0; s < E; ++s) Ps[s] = lds[(b ^ fpfft::lds_mask(s * TPB)) + s * TPB];
  } else {
#pragma unroll
    for (int s = 0; s < E; ++s) Ps[s] = lds[fpfft::lds_pad(tid + s * TPB)];
  }
  __syncthreads();
  const cpx<T> wh = wt[NT / 2];
#pragma unroll
  for (int s = 0; s < E; ++s) {
    const int s2 = (s + E / 2) % E;
    const cpx<T> q = mk<T>(Is[s2] + Ps[s2].x, Ps[s2].y);
    KR[s] = Ps[s].y - (wh * q).y;
  }
}

}  // namespace fpraman
