// C-ABI entry points of the gfx950 propagator library (declared in
// include/rapport_amd.h) and the kernel instantiations they dispatch to.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>

#include "../../include/rapport_amd.h"
#include "common.hpp"
#include "fft.hpp"
#include "gemm_f64.hpp"
#include "gemm_mixed.hpp"
#include "nl.hpp"
#include "nl_f32.hpp"
#include "plasma.hpp"

#define FP_VERSION 100

namespace {

thread_local std::string g_err;

int fail(int code, const char* what) {
  g_err = what;
  return code;
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return (int)e;
  }
  return FP_OK;
}

inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

// ------------------------------------------------------------------ FFT rows
template <typename T, int NT, int TPB, bool INV>
__global__ __launch_bounds__(TPB) void fft_rows_kernel(const T* in, T* out, const T* __restrict__ tw) {
  __shared__ cpx<T> lds[fpfft::LdsSize<NT>::value];
  constexpr int E = NT / TPB;
  cpx<T> v[E];
  const size_t row = blockIdx.x;
  const cpx<T>* src = reinterpret_cast<const cpx<T>*>(in) + row * NT;
#pragma unroll
  for (int i = 0; i < E; ++i) v[i] = src[threadIdx.x + i * TPB];
  fpfft::row_fft<T, NT, TPB, INV>(v, lds, reinterpret_cast<const cpx<T>*>(tw));
  if (INV) {
    const T s = T(1) / (T)NT;
#pragma unroll
    for (int i = 0; i < E; ++i) v[i] = s * v[i];
  }
  cpx<T>* dst = reinterpret_cast<cpx<T>*>(out) + row * NT;
#pragma unroll
  for (int i = 0; i < E; ++i) dst[threadIdx.x + i * TPB] = v[i];
}

template <typename T, int NT, int TPB>
int launch_fft_rows(const T* in, T* out, int Nr, int inverse, const T* tw, hipStream_t st) {
  if (inverse)
    hipLaunchKernelGGL((fft_rows_kernel<T, NT, TPB, true>), dim3(Nr), dim3(TPB), 0, st, in, out, tw);
  else
    hipLaunchKernelGGL((fft_rows_kernel<T, NT, TPB, false>), dim3(Nr), dim3(TPB), 0, st, in, out, tw);
  return check_launch("fft_rows");
}

template <typename T>
int dispatch_fft_rows(const T* in, T* out, int Nr, int Nt, int inverse, const T* tw, hipStream_t st) {
  switch (Nt) {
    case 512: return launch_fft_rows<T, 512, 64>(in, out, Nr, inverse, tw, st);
    case 1024: return launch_fft_rows<T, 1024, 128>(in, out, Nr, inverse, tw, st);
    case 2048: return launch_fft_rows<T, 2048, 256>(in, out, Nr, inverse, tw, st);
    case 4096: return launch_fft_rows<T, 4096, 256>(in, out, Nr, inverse, tw, st);
    default: return fail(FP_E_SHAPE, "fft rows: Nt must be 512, 1024, 2048 or 4096");
  }
}

__global__ void absmax_kernel(const float* __restrict__ x, long long n, unsigned* slot) {
  __shared__ float sred[4];
  float m = 0.0f;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    m = fmaxf(m, fabsf(x[i]));
  for (int d = 32; d >= 1; d >>= 1) m = fmaxf(m, __shfl_xor(m, d, 64));
  if ((threadIdx.x & 63) == 0) sred[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; ++w) m = fmaxf(m, sred[w]);
    atomicMax(slot, __float_as_uint(m));
  }
}

// ------------------------------------------------------------------ plasma
constexpr int cmax(int a, int b) { return a > b ? a : b; }

template <typename T, int NT, int TPB, bool STE, bool SPEC>
__global__ __launch_bounds__(TPB) void plasma_kernel(const T* __restrict__ src,
                                                      double* __restrict__ rho,
                                                      double* __restrict__ rhos,
                                                      T* __restrict__ u_out,
                                                      fp_plasma_params p,
                                                      const double* __restrict__ Wg,
                                                      const double* __restrict__ Ws,
                                                      const T* __restrict__ tw, int* status,
                                                      unsigned* clear_slots) {
  constexpr int PL = fpplasma::PlasmaLds<NT, TPB>::bytes;
  constexpr int FB = fpfft::LdsSize<NT>::value * 2 * sizeof(T);
  __shared__ __attribute__((aligned(16))) char smem[cmax(PL, FB)];
  double* sI = reinterpret_cast<double*>(smem);
  const size_t row = blockIdx.x;
  if (clear_slots && blockIdx.x == 0 && threadIdx.x < 3) clear_slots[threadIdx.x] = 0u;
  if (SPEC) {
    constexpr int E = NT / TPB;
    cpx<T> v[E];
    const cpx<T>* sp = reinterpret_cast<const cpx<T>*>(src) + row * NT;
#pragma unroll
    for (int i = 0; i < E; ++i) v[i] = sp[threadIdx.x + i * TPB];
    fpfft::row_fft<T, NT, TPB, true>(v, reinterpret_cast<cpx<T>*>(smem),
                                     reinterpret_cast<const cpx<T>*>(tw));
    const T s = T(1) / (T)NT;
#pragma unroll
    for (int i = 0; i < E; ++i) v[i] = s * v[i];
    if (u_out) {
      cpx<T>* uo = reinterpret_cast<cpx<T>*>(u_out) + row * NT;
#pragma unroll
      for (int i = 0; i < E; ++i) uo[threadIdx.x + i * TPB] = v[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const double x = (double)v[i].x, y = (double)v[i].y;
      sI[threadIdx.x + i * TPB] = (x * x + y * y) * p.field_to_I;
    }
  } else {
    const cpx<T>* ur = reinterpret_cast<const cpx<T>*>(src) + row * NT;
    for (int k = threadIdx.x; k < NT; k += TPB) {
      const cpx<T> z = ur[k];
      const double x = (double)z.x, y = (double)z.y;
      sI[k] = (x * x + y * y) * p.field_to_I;
    }
  }
  __syncthreads();
  fpplasma::plasma_row<NT, TPB, STE>(sI, p, Wg, Ws, rho + row * NT,
                                     rhos ? rhos + row * NT : nullptr, status);
}

template <typename T, int NT, bool SPEC>
int launch_plasma(const T* src, double* rho, double* rhos, T* u_out, const fp_plasma_params& p,
                  const double* Wg, const double* Ws, const T* tw, int* status,
                  unsigned* clear_slots, hipStream_t st) {
  constexpr int TPB = 256;
  if (p.enable_ste)
    hipLaunchKernelGGL((plasma_kernel<T, NT, TPB, true, SPEC>), dim3(p.Nr), dim3(TPB), 0, st, src,
                       rho, rhos, u_out, p, Wg, Ws, tw, status, clear_slots);
  else
    hipLaunchKernelGGL((plasma_kernel<T, NT, TPB, false, SPEC>), dim3(p.Nr), dim3(TPB), 0, st, src,
                       rho, rhos, u_out, p, Wg, Ws, tw, status, clear_slots);
  return check_launch("plasma_scan");
}

template <typename T, bool SPEC>
int dispatch_plasma(const T* src, double* rho, double* rhos, T* u_out, const fp_plasma_params* p,
                    const double* Wg, const double* Ws, const T* tw, int* status,
                    unsigned* clear_slots, hipStream_t st) {
  if (!src || !rho || !p || !Wg) return fail(FP_E_ARG, "plasma scan: null pointer");
  if (SPEC && !tw) return fail(FP_E_ARG, "plasma scan: spectral input needs twiddles");
  if (p->enable_ste && (!rhos || !Ws)) return fail(FP_E_ARG, "plasma scan: STE needs rho_s, Ws_lut");
  if (p->Nr <= 0 || p->table_size < 2) return fail(FP_E_ARG, "plasma scan: bad Nr/table_size");
  switch (p->Nt) {
    case 512: return launch_plasma<T, 512, SPEC>(src, rho, rhos, u_out, *p, Wg, Ws, tw, status, clear_slots, st);
    case 1024: return launch_plasma<T, 1024, SPEC>(src, rho, rhos, u_out, *p, Wg, Ws, tw, status, clear_slots, st);
    case 2048: return launch_plasma<T, 2048, SPEC>(src, rho, rhos, u_out, *p, Wg, Ws, tw, status, clear_slots, st);
    case 4096: return launch_plasma<T, 4096, SPEC>(src, rho, rhos, u_out, *p, Wg, Ws, tw, status, clear_slots, st);
    default: return fail(FP_E_SHAPE, "plasma scan: Nt must be 512, 1024, 2048 or 4096");
  }
}

// ------------------------------------------------------------------ nonlinear
template <int NT> struct NLT { static constexpr int TPB = NT >= 4096 ? NT / 8 : NT / 4; };
template <int NT> struct NLT32 { static constexpr int TPB = NT / 8; };

}  // namespace

extern "C" {

int fp_version(void) { return FP_VERSION; }
const char* fp_last_error(void) { return g_err.c_str(); }

int fp_plasma_scan(const double* u, double* rho, double* rho_s, const fp_plasma_params* p,
                   const double* W_lut, const double* Ws_lut, int* status, void* stream) {
  return dispatch_plasma<double, false>(u, rho, rho_s, nullptr, p, W_lut, Ws_lut, nullptr, status,
                                        nullptr, S(stream));
}

int fp_plasma_scan_spec(const double* spec, double* rho, double* rho_s, double* u_out,
                        const fp_plasma_params* p, const double* W_lut, const double* Ws_lut,
                        const double* tw, int* status, void* stream) {
  return dispatch_plasma<double, true>(spec, rho, rho_s, u_out, p, W_lut, Ws_lut, tw, status,
                                       nullptr, S(stream));
}

int fp_fft_rows(const double* in, double* out, int Nr, int Nt, int inverse, const double* tw,
                void* stream) {
  if (!in || !out || !tw || Nr <= 0) return fail(FP_E_ARG, "fp_fft_rows: bad argument");
  return dispatch_fft_rows<double>(in, out, Nr, Nt, inverse, tw, S(stream));
}

int fp_hankel_gemm(const double* Yt, const double* B, double* C, int Nr, int Nt, double alpha,
                   const double* rowscale, const double* phase, void* stream) {
  if (!Yt || !B || !C || Nr <= 0 || Nt <= 0) return fail(FP_E_ARG, "fp_hankel_gemm: bad argument");
  if (B == C) return fail(FP_E_ARG, "fp_hankel_gemm: C must not alias B");
  const int Ncols = 2 * Nt;
  const int nb = ((Nr + fpgemm::BM - 1) / fpgemm::BM) * ((Ncols + fpgemm::BN - 1) / fpgemm::BN);
  hipLaunchKernelGGL(fpgemm::hankel_gemm_f64, dim3(nb), dim3(fpgemm::NTHR), 0, S(stream), Yt, B, C,
                     Nr, Nr, Ncols, alpha, rowscale, phase);
  return check_launch("hankel_gemm");
}

int fp_half_linear(const double* u, double* out, int Nr, int Nt, const double* Yt,
                   const double* phase_half, double Rk, double iRk, const double* tw,
                   double* work, void* stream) {
  if (!u || !out || !Yt || !phase_half || !tw || !work) return fail(FP_E_ARG, "fp_half_linear: null pointer");
  double* w0 = work;
  double* w1 = work + (size_t)2 * Nr * Nt;
  int rc;
  if ((rc = fp_fft_rows(u, w0, Nr, Nt, 0, tw, stream))) return rc;
  if ((rc = fp_hankel_gemm(Yt, w0, w1, Nr, Nt, Rk, nullptr, phase_half, stream))) return rc;
  if ((rc = fp_hankel_gemm(Yt, w1, w0, Nr, Nt, iRk, nullptr, nullptr, stream))) return rc;
  return fp_fft_rows(w0, out, Nr, Nt, 1, tw, stream);
}

int fp_nl_split(const double* u, const double* rho, const double* rho_s, double* rhs,
                double* alpha, const fp_nl_params* p, const fp_nl_tables* t, void* stream) {
  if (!u || !rho || !rhs || !p || !t) return fail(FP_E_ARG, "fp_nl_split: null pointer");
  hipStream_t st = S(stream);
#define FP_SPLIT(NT)                                                                            \
  if (rho_s)                                                                                    \
    hipLaunchKernelGGL((fpnl::split_kernel<NT, NLT<NT>::TPB, true>), dim3(p->Nr),               \
                       dim3(NLT<NT>::TPB), 0, st, u, rho, rho_s, rhs, alpha, *p, *t);           \
  else                                                                                          \
    hipLaunchKernelGGL((fpnl::split_kernel<NT, NLT<NT>::TPB, false>), dim3(p->Nr),              \
                       dim3(NLT<NT>::TPB), 0, st, u, rho, rho_s, rhs, alpha, *p, *t);           \
  return check_launch("nl_split");
  switch (p->Nt) {
    case 512: { FP_SPLIT(512) }
    case 1024: { FP_SPLIT(1024) }
    case 2048: { FP_SPLIT(2048) }
    case 4096: { FP_SPLIT(4096) }
    default: return fail(FP_E_SHAPE, "fp_nl_split: Nt must be 512, 1024, 2048 or 4096");
  }
#undef FP_SPLIT
}

int fp_nl_step(const double* in, double* out, const double* rho, const double* rho_s,
               int in_spectral, int out_spectral, const fp_nl_params* p, const fp_nl_tables* t,
               void* stream) {
  if (!in || !out || !rho || !p || !t) return fail(FP_E_ARG, "fp_nl_step: null pointer");
  hipStream_t st = S(stream);
#define FP_STEP(NT)                                                                             \
  if (rho_s)                                                                                    \
    hipLaunchKernelGGL((fpnl::nl_step_kernel<NT, NLT<NT>::TPB, true>), dim3(p->Nr),             \
                       dim3(NLT<NT>::TPB), 0, st, in, out, rho, rho_s, in_spectral,             \
                       out_spectral, *p, *t);                                                   \
  else                                                                                          \
    hipLaunchKernelGGL((fpnl::nl_step_kernel<NT, NLT<NT>::TPB, false>), dim3(p->Nr),            \
                       dim3(NLT<NT>::TPB), 0, st, in, out, rho, rho_s, in_spectral,             \
                       out_spectral, *p, *t);                                                   \
  return check_launch("nl_step");
  switch (p->Nt) {
    case 512: { FP_STEP(512) }
    case 1024: { FP_STEP(1024) }
    case 2048: { FP_STEP(2048) }
    case 4096: { FP_STEP(4096) }
    default: return fail(FP_E_SHAPE, "fp_nl_step: Nt must be 512, 1024, 2048 or 4096");
  }
#undef FP_STEP
}

int fp_record_rows(const double* u, const double* rho, const double* rho_s, double* rows_out,
                   const fp_nl_params* p, const fp_nl_tables* t, double dt, int nbins,
                   const int* bin_lo, const int* bin_hi, double* bins, void* stream) {
  if (!u || !rho || !rows_out || !p || !t) return fail(FP_E_ARG, "fp_record_rows: null pointer");
  if (nbins > 0 && (!bin_lo || !bin_hi || !bins)) return fail(FP_E_ARG, "fp_record_rows: bins");
  hipLaunchKernelGGL((fpnl::record_kernel<256>), dim3(p->Nr), dim3(256), 0, S(stream), u, rho,
                     rho_s, rows_out, *p, *t, dt, nbins, bin_lo, bin_hi, bins);
  return check_launch("record_rows");
}

// ------------------------------------------------------------------ mixed precision
int fp_absmax_f32(const float* x, long long n, unsigned* slot, void* stream) {
  if (!x || !slot || n < 0) return fail(FP_E_ARG, "fp_absmax_f32: bad argument");
  long long blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(absmax_kernel, dim3((unsigned)blocks), dim3(256), 0, S(stream), x, n, slot);
  return check_launch("absmax");
}

int fp_fft_rows_c64(const float* in, float* out, int Nr, int Nt, int inverse, const float* tw,
                    void* stream) {
  if (!in || !out || !tw || Nr <= 0) return fail(FP_E_ARG, "fp_fft_rows_c64: bad argument");
  return dispatch_fft_rows<float>(in, out, Nr, Nt, inverse, tw, S(stream));
}

int fp_plasma_scan_spec_c64(const float* spec, double* rho, double* rho_s, float* u_out,
                            const fp_plasma_params* p, const double* W_lut, const double* Ws_lut,
                            const float* tw, unsigned* clear_slots, int* status, void* stream) {
  return dispatch_plasma<float, true>(spec, rho, rho_s, u_out, p, W_lut, Ws_lut, tw, status,
                                      clear_slots, S(stream));
}

int fp_hankel_gemm_mixed(const void* Yh, const void* Yl, int y_scale_exp, const float* B,
                         float* C, int Nr, int Nt, float alpha, const float* rowscale,
                         const float* phase, const unsigned* bmax_in, unsigned* cmax_out,
                         unsigned* reset_slot, void* stream) {
  if (!Yh || !Yl || !B || !C || !bmax_in) return fail(FP_E_ARG, "fp_hankel_gemm_mixed: null pointer");
  if (B == C) return fail(FP_E_ARG, "fp_hankel_gemm_mixed: C must not alias B");
  const int Ncols = 2 * Nt;
  if (Nr <= 0 || Nr % fpgemm16::BM || Ncols % fpgemm16::BN)
    return fail(FP_E_SHAPE, "fp_hankel_gemm_mixed: needs Nr % 128 == 0 and 2 Nt % 128 == 0");
  const int nb = (Nr / fpgemm16::BM) * (Ncols / fpgemm16::BN);
  hipLaunchKernelGGL(fpgemm16::hankel_gemm_f16x3, dim3(nb), dim3(fpgemm16::NTHR), 0, S(stream),
                     reinterpret_cast<const fpgemm16::h8*>(Yh),
                     reinterpret_cast<const fpgemm16::h8*>(Yl), y_scale_exp, B, C, Nr, Nr, Ncols,
                     alpha, rowscale, phase, bmax_in, cmax_out, reset_slot);
  return check_launch("hankel_gemm_mixed");
}

int fp_nl_split_c64(const float* u, const double* rho, const double* rho_s, float* rhs,
                    float* alpha, const fp_nl_params* p, const fp_nl_tables_f32* t,
                    void* stream) {
  if (!u || !rho || !rhs || !p || !t) return fail(FP_E_ARG, "fp_nl_split_c64: null pointer");
  hipStream_t st = S(stream);
#define FP_SPLIT32(NT)                                                                           \
  if (rho_s)                                                                                     \
    hipLaunchKernelGGL((fpnl32::split_kernel<NT, NLT32<NT>::TPB, true>), dim3(p->Nr),            \
                       dim3(NLT32<NT>::TPB), 0, st, u, rho, rho_s, rhs, alpha, *p, *t);          \
  else                                                                                           \
    hipLaunchKernelGGL((fpnl32::split_kernel<NT, NLT32<NT>::TPB, false>), dim3(p->Nr),           \
                       dim3(NLT32<NT>::TPB), 0, st, u, rho, rho_s, rhs, alpha, *p, *t);          \
  return check_launch("nl_split_c64");
  switch (p->Nt) {
    case 512: { FP_SPLIT32(512) }
    case 1024: { FP_SPLIT32(1024) }
    case 2048: { FP_SPLIT32(2048) }
    case 4096: { FP_SPLIT32(4096) }
    default: return fail(FP_E_SHAPE, "fp_nl_split_c64: Nt must be 512, 1024, 2048 or 4096");
  }
#undef FP_SPLIT32
}

int fp_nl_step_c64(const float* in, float* out, const double* rho, const double* rho_s,
                   int in_spectral, int out_spectral, const fp_nl_params* p,
                   const fp_nl_tables_f32* t, unsigned* cmax_out, void* stream) {
  if (!in || !out || !rho || !p || !t) return fail(FP_E_ARG, "fp_nl_step_c64: null pointer");
  hipStream_t st = S(stream);
#define FP_STEP32(NT)                                                                            \
  if (rho_s)                                                                                     \
    hipLaunchKernelGGL((fpnl32::nl_step_kernel<NT, NLT32<NT>::TPB, true>), dim3(p->Nr),          \
                       dim3(NLT32<NT>::TPB), 0, st, in, out, rho, rho_s, in_spectral,            \
                       out_spectral, *p, *t, cmax_out);                                          \
  else                                                                                           \
    hipLaunchKernelGGL((fpnl32::nl_step_kernel<NT, NLT32<NT>::TPB, false>), dim3(p->Nr),         \
                       dim3(NLT32<NT>::TPB), 0, st, in, out, rho, rho_s, in_spectral,            \
                       out_spectral, *p, *t, cmax_out);                                          \
  return check_launch("nl_step_c64");
  switch (p->Nt) {
    case 512: { FP_STEP32(512) }
    case 1024: { FP_STEP32(1024) }
    case 2048: { FP_STEP32(2048) }
    case 4096: { FP_STEP32(4096) }
    default: return fail(FP_E_SHAPE, "fp_nl_step_c64: Nt must be 512, 1024, 2048 or 4096");
  }
#undef FP_STEP32
}

}  // extern "C"
