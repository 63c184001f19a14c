// C-ABI entry points of the gfx950 propagator library (declared in
// include/rapport_amd.h) and the kernel instantiations they dispatch to.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>

#include "../../include/rapport_amd.h"
#include "common.hpp"
#include "fft.hpp"
#include "gemm_f64.hpp"
#include "nl.hpp"
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
template <int NT, int TPB, bool INV>
__global__ __launch_bounds__(TPB) void fft_rows_kernel(const double* __restrict__ in,
                                                        double* __restrict__ out,
                                                        const double* __restrict__ tw) {
  __shared__ cpx<double> lds[fpfft::LdsSize<NT>::value];
  constexpr int E = NT / TPB;
  cpx<double> v[E];
  const size_t row = blockIdx.x;
  fpnl::load_row<NT, TPB>(in, row, v);
  fpfft::row_fft<double, NT, TPB, INV>(v, lds, reinterpret_cast<const cpx<double>*>(tw));
  if (INV) {
    const double s = 1.0 / (double)NT;
#pragma unroll
    for (int i = 0; i < E; ++i) v[i] = s * v[i];
  }
  fpnl::store_row<NT, TPB>(out, row, v);
}

template <int NT, int TPB>
int launch_fft_rows(const double* in, double* out, int Nr, int inverse, const double* tw,
                    hipStream_t st) {
  if (inverse)
    hipLaunchKernelGGL((fft_rows_kernel<NT, TPB, true>), dim3(Nr), dim3(TPB), 0, st, in, out, tw);
  else
    hipLaunchKernelGGL((fft_rows_kernel<NT, TPB, false>), dim3(Nr), dim3(TPB), 0, st, in, out, tw);
  return check_launch("fft_rows");
}

// ------------------------------------------------------------------ plasma
constexpr int cmax(int a, int b) { return a > b ? a : b; }

template <int NT, int TPB, bool STE, bool SPEC>
__global__ __launch_bounds__(TPB) void plasma_kernel(const double* __restrict__ src,
                                                      double* __restrict__ rho,
                                                      double* __restrict__ rhos,
                                                      double* __restrict__ u_out,
                                                      fp_plasma_params p,
                                                      const double* __restrict__ Wg,
                                                      const double* __restrict__ Ws,
                                                      const double* __restrict__ tw, int* status) {
  constexpr int PL = fpplasma::PlasmaLds<NT, TPB>::bytes;
  constexpr int FB = fpfft::LdsSize<NT>::value * 16;
  __shared__ __attribute__((aligned(16))) char smem[cmax(PL, FB)];
  double* sI = reinterpret_cast<double*>(smem);
  const size_t row = blockIdx.x;
  if (SPEC) {
    constexpr int E = NT / TPB;
    cpx<double> v[E];
    fpnl::load_row<NT, TPB>(src, row, v);
    fpfft::row_fft<double, NT, TPB, true>(v, reinterpret_cast<cpx<double>*>(smem),
                                          reinterpret_cast<const cpx<double>*>(tw));
    const double s = 1.0 / (double)NT;
#pragma unroll
    for (int i = 0; i < E; ++i) v[i] = s * v[i];
    if (u_out) fpnl::store_row<NT, TPB>(u_out, row, v);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < E; ++i)
      sI[threadIdx.x + i * TPB] = (v[i].x * v[i].x + v[i].y * v[i].y) * p.field_to_I;
  } else {
    const cpx<double>* ur = reinterpret_cast<const cpx<double>*>(src) + row * NT;
    for (int k = threadIdx.x; k < NT; k += TPB) {
      const cpx<double> z = ur[k];
      sI[k] = (z.x * z.x + z.y * z.y) * p.field_to_I;
    }
  }
  __syncthreads();
  fpplasma::plasma_row<NT, TPB, STE>(sI, p, Wg, Ws, rho + row * NT,
                                     rhos ? rhos + row * NT : nullptr, status);
}

template <int NT, bool SPEC>
int launch_plasma(const double* src, double* rho, double* rhos, double* u_out,
                  const fp_plasma_params& p, const double* Wg, const double* Ws,
                  const double* tw, int* status, hipStream_t st) {
  constexpr int TPB = 256;
  if (p.enable_ste)
    hipLaunchKernelGGL((plasma_kernel<NT, TPB, true, SPEC>), dim3(p.Nr), dim3(TPB), 0, st, src,
                       rho, rhos, u_out, p, Wg, Ws, tw, status);
  else
    hipLaunchKernelGGL((plasma_kernel<NT, TPB, false, SPEC>), dim3(p.Nr), dim3(TPB), 0, st, src,
                       rho, rhos, u_out, p, Wg, Ws, tw, status);
  return check_launch("plasma_scan");
}

// ------------------------------------------------------------------ nonlinear
template <int NT> struct NLT { static constexpr int TPB = NT >= 4096 ? NT / 8 : NT / 4; };

}  // namespace

extern "C" {

int fp_version(void) { return FP_VERSION; }
const char* fp_last_error(void) { return g_err.c_str(); }

int fp_plasma_scan(const double* u, double* rho, double* rho_s, const fp_plasma_params* p,
                   const double* W_lut, const double* Ws_lut, int* status, void* stream) {
  if (!u || !rho || !p || !W_lut) return fail(FP_E_ARG, "fp_plasma_scan: null pointer");
  if (p->enable_ste && (!rho_s || !Ws_lut)) return fail(FP_E_ARG, "fp_plasma_scan: STE needs rho_s, Ws_lut");
  if (p->Nr <= 0 || p->table_size < 2) return fail(FP_E_ARG, "fp_plasma_scan: bad Nr/table_size");
  switch (p->Nt) {
    case 512: return launch_plasma<512, false>(u, rho, rho_s, nullptr, *p, W_lut, Ws_lut, nullptr, status, S(stream));
    case 1024: return launch_plasma<1024, false>(u, rho, rho_s, nullptr, *p, W_lut, Ws_lut, nullptr, status, S(stream));
    case 2048: return launch_plasma<2048, false>(u, rho, rho_s, nullptr, *p, W_lut, Ws_lut, nullptr, status, S(stream));
    case 4096: return launch_plasma<4096, false>(u, rho, rho_s, nullptr, *p, W_lut, Ws_lut, nullptr, status, S(stream));
    default: return fail(FP_E_SHAPE, "fp_plasma_scan: Nt must be 512, 1024, 2048 or 4096");
  }
}

int fp_plasma_scan_spec(const double* spec, double* rho, double* rho_s, double* u_out,
                        const fp_plasma_params* p, const double* W_lut, const double* Ws_lut,
                        const double* tw, int* status, void* stream) {
  if (!spec || !rho || !p || !W_lut || !tw) return fail(FP_E_ARG, "fp_plasma_scan_spec: null pointer");
  if (p->enable_ste && (!rho_s || !Ws_lut)) return fail(FP_E_ARG, "fp_plasma_scan_spec: STE needs rho_s, Ws_lut");
  if (p->Nr <= 0 || p->table_size < 2) return fail(FP_E_ARG, "fp_plasma_scan_spec: bad Nr/table_size");
  switch (p->Nt) {
    case 512: return launch_plasma<512, true>(spec, rho, rho_s, u_out, *p, W_lut, Ws_lut, tw, status, S(stream));
    case 1024: return launch_plasma<1024, true>(spec, rho, rho_s, u_out, *p, W_lut, Ws_lut, tw, status, S(stream));
    case 2048: return launch_plasma<2048, true>(spec, rho, rho_s, u_out, *p, W_lut, Ws_lut, tw, status, S(stream));
    case 4096: return launch_plasma<4096, true>(spec, rho, rho_s, u_out, *p, W_lut, Ws_lut, tw, status, S(stream));
    default: return fail(FP_E_SHAPE, "fp_plasma_scan_spec: Nt must be 512, 1024, 2048 or 4096");
  }
}

int fp_fft_rows(const double* in, double* out, int Nr, int Nt, int inverse, const double* tw,
                void* stream) {
  if (!in || !out || !tw || Nr <= 0) return fail(FP_E_ARG, "fp_fft_rows: bad argument");
  switch (Nt) {
    case 512: return launch_fft_rows<512, 64>(in, out, Nr, inverse, tw, S(stream));
    case 1024: return launch_fft_rows<1024, 128>(in, out, Nr, inverse, tw, S(stream));
    case 2048: return launch_fft_rows<2048, 256>(in, out, Nr, inverse, tw, S(stream));
    case 4096: return launch_fft_rows<4096, 256>(in, out, Nr, inverse, tw, S(stream));
    default: return fail(FP_E_SHAPE, "fp_fft_rows: Nt must be 512, 1024, 2048 or 4096");
  }
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

}  // extern "C"
