// C-ABI entry points of the gfx950 propagator library (declared in
// include/rapport_amd.h) and the kernel instantiations they dispatch to.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rapport_amd.h"
#include "adjoint.hpp"
#include "common.hpp"
#include "fft.hpp"
#include "gemm_f64.hpp"
#include "gemm_mixed.hpp"
#include "nl.hpp"
#include "nl_f32.hpp"
#include "plasma.hpp"
#include "rate0d.hpp"
#include "record.hpp"
#include "shock1d.hpp"

#define FP_VERSION 100

// 16-row blocks per wave whose phase values the single-instance mixed GEMM loads during
// its last K slabs (gemm_mixed.hpp); the rest are loaded at the start of its epilogue.
#ifndef FP_MIXED_GEMM_NPF
#define FP_MIXED_GEMM_NPF 1
#endif

namespace {

thread_local std::string g_err;

int fail(int code, const char* what) {
  g_err = what;
  return code;
}

int device_cus();   // compute units of the current device (defined below)

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

// max |x_i| into *slot (float bits, atomicMax). Few blocks (one atomic each: the
// atomics on one word serialise) and 16-byte loads where the length allows.
__global__ void absmax_kernel(const float* __restrict__ x, long long n, unsigned* slot) {
  __shared__ float sred[4];
  float m = 0.0f;
  const long long stride = (long long)gridDim.x * blockDim.x;
  const long long n4 = n / 4;
  const float4* x4 = reinterpret_cast<const float4*>(x);
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) {
    const float4 v = x4[i];
    m = fmaxf(m, fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w))));
  }
  for (long long i = 4 * n4 + (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
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

template <typename T, int NT, int TPB, bool STE>
struct PlasmaSmem {
  static constexpr int bytes = cmax(fpplasma::PlasmaLds<NT, TPB, STE>::bytes,
                                    fpfft::LdsSize<NT>::value * 2 * (int)sizeof(T));
};

// Minimum waves per SIMD the carrier-scan kernel is built for (second launch-bounds
// argument; it caps VGPRs at 512 / value). A 256-thread row is one wave per SIMD, so this is
// also the number of row blocks per CU. A complex64 row without STE takes 37 KB of LDS, so up
// to four rows fit in one CU; three keep the kernel at 164 VGPRs without spills (four would
// need 128 and spill). The other variants take 64-70 KB of LDS or more than 256 VGPRs and
// keep one.
#ifndef FP_PLASMA_OCC_C64
#define FP_PLASMA_OCC_C64 3
#endif
template <typename T, bool STE>
struct PlasmaOcc {
  static constexpr int value = (!STE && sizeof(T) == 4) ? FP_PLASMA_OCC_C64 : 1;
};

// One row: (optional inverse FFT of a spectral row) -> intensity -> carrier scan.
template <typename T, int NT, int TPB, bool STE, bool SPEC>
FP_DEV void plasma_row_body(char* smem, size_t row, const T* __restrict__ src,
                            double* __restrict__ rho, double* __restrict__ rhos,
                            T* __restrict__ u_out, const fp_plasma_params& p,
                            const double* __restrict__ Wg, const double* __restrict__ Ws,
                            const T* __restrict__ tw, int* status) {
  double* sI = reinterpret_cast<double*>(smem);
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

template <typename T, int NT, int TPB, bool STE, bool SPEC>
__global__ __launch_bounds__(TPB, (PlasmaOcc<T, STE>::value)) void plasma_kernel(const T* __restrict__ src,
                                                      double* __restrict__ rho,
                                                      double* __restrict__ rhos,
                                                      T* __restrict__ u_out,
                                                      fp_plasma_params p,
                                                      const double* __restrict__ Wg,
                                                      const double* __restrict__ Ws,
                                                      const T* __restrict__ tw, int* status,
                                                      unsigned* clear_slots) {
  __shared__ __attribute__((aligned(16))) char smem[PlasmaSmem<T, NT, TPB, STE>::bytes];
  if (clear_slots && blockIdx.x == 0 && threadIdx.x < 3) clear_slots[threadIdx.x] = 0u;
  plasma_row_body<T, NT, TPB, STE, SPEC>(smem, blockIdx.x, src, rho, rhos, u_out, p, Wg, Ws, tw,
                                         status);
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
// Samples per thread of the nonlinear-step rows, from measured step times: FP64 rows take 8
// from 2048 samples on (512 x 2048: 0.471 vs 0.529 ms per step with 4) and 4 below (more
// waves for the small grids); FP32 rows of 4096 samples take 16 (256 threads, three radix-16
// passes per transform, 80 KiB of LDS, two blocks per CU: 12 % less time than 8 samples at
// 1024 x 4096), 8 at 2048 and 4 below (128 x 512: 0.0876 vs 0.0935 ms per mixed step with 8).
template <int NT> struct NLT { static constexpr int TPB = NT >= 2048 ? NT / 8 : NT / 4; };
template <int NT> struct NLT32 {
  static constexpr int TPB = NT >= 4096 ? NT / 16 : (NT >= 2048 ? NT / 8 : NT / 4);
};

// ------------------------------------------------------------------ batched form
// Work items (rows of the row kernels, output tiles of the GEMM) of every instance
// of a group are numbered consecutively: instance g.idx[j] owns items
// [g.begin[j], g.begin[j+1]). A block maps its item to (instance, local item) by a
// binary search over the group (uniform per block, so scalar arithmetic); with a
// persistent grid (gridDim.x < items) each block loops over items.
constexpr int kGroupMax = 128;
// The descriptor is shared between host and device code: same layout on both sides.
static_assert(sizeof(fp_batch_instance) == 448 && offsetof(fp_batch_instance, pp) == 136 &&
                  offsetof(fp_batch_instance, np) == 240 && offsetof(fp_batch_instance, nt) == 368,
              "fp_batch_instance layout");
struct BatchGroup {
  int n;
  int idx[kGroupMax];
  int begin[kGroupMax + 1];
};

FP_DEV int group_find(const BatchGroup& g, int item) {
  int lo = 0, hi = g.n;   // g.begin[lo] <= item < g.begin[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (g.begin[mid] <= item) lo = mid; else hi = mid;
  }
  return lo;
}

// A device pointer stored in an instance descriptor, read as a global-memory pointer. A
// pointer the kernel loads from memory is generic, so the accesses through it compile to
// flat instructions: 64-bit VGPR addresses, and each flat access also counts against the LDS
// wait counter, so every wait for an LDS read then waits for the outstanding flat loads too.
// Read through a global-address-space type, the uses become global instructions (scalar base
// plus 32-bit offset), as for the single-instance kernels' pointer arguments.
template <typename T>
FP_DEV T* gp(T* const& field) {
  using GT = __attribute__((address_space(1))) T;
  return (T*)*reinterpret_cast<GT* const*>(reinterpret_cast<const char*>(&field));
}

// The float32 nonlinear-step tables of a descriptor with every pointer read by gp().
FP_DEV fp_nl_tables_f32 gtables(const fp_nl_tables_f32& t) {
  fp_nl_tables_f32 r;
  r.tw = gp(t.tw);
  r.mA = gp(t.mA);
  r.mB = gp(t.mB);
  r.mC = gp(t.mC);
  r.Rf = gp(t.Rf);
  r.knots = gp(t.knots);
  r.inv_h = gp(t.inv_h);
  r.pcoef_s = gp(t.pcoef_s);
  r.rw = gp(t.rw);
  return r;
}

// PERS: a persistent grid (blocks loop over the items); otherwise one block per item, with
// no loop state to carry (the 4096-sample scan then fits the single-instance kernel's
// register budget: three blocks per CU, no spills).
template <int NT, int TPB, bool STE, bool PERS>
__global__ __launch_bounds__(TPB, (PlasmaOcc<float, STE>::value)) void batch_plasma_kernel(const fp_batch_instance* __restrict__ insts,
                                                           BatchGroup g) {
  __shared__ __attribute__((aligned(16))) char smem[PlasmaSmem<float, NT, TPB, STE>::bytes];
  auto one = [&](int item) {
    const int j = group_find(g, item);
    const fp_batch_instance& I = insts[__builtin_amdgcn_readfirstlane(g.idx[j])];
    const int row = item - g.begin[j];
    // every descriptor field is read before the first store (the slot clear): reads that
    // follow a store through a pointer could alias the descriptors, so the compiler would
    // make them vector loads, holding each pointer in VGPRs (74 spilled VGPRs at 4096)
    const fp_plasma_params pp = I.pp;
    const float* spec = gp(I.spec);
    double *rho = gp(I.rho), *rho_s = STE ? gp(I.rho_s) : nullptr;
    const double *W = gp(I.W_lut), *Ws = gp(I.Ws_lut);
    const float* tw = gp(I.tw);
    int* status = gp(I.status);
    unsigned* slots = gp(I.slots);
    if (row == 0 && threadIdx.x < 3) slots[1 + threadIdx.x] = 0u;
    plasma_row_body<float, NT, TPB, STE, true>(smem, row, spec, rho, rho_s, nullptr, pp, W, Ws,
                                               tw, status);
  };
  if constexpr (PERS) {
    const int total = g.begin[g.n];
    for (int item = blockIdx.x; item < total; item += gridDim.x) {
      one(item);
      __syncthreads();
    }
  } else {
    one(blockIdx.x);
  }
}

template <int NT, int TPB, bool RS, bool PERS>
__global__ __launch_bounds__(TPB, (TPB >= 512 ? 4 : 2)) void batch_nl_kernel(
    const fp_batch_instance* __restrict__ insts, BatchGroup g) {
  __shared__ fpnl32::StepSmem<NT> sm;
  auto one = [&](int item) {
    const int j = group_find(g, item);
    const fp_batch_instance& I = insts[g.idx[j]];
    const int row = item - g.begin[j];
    float* w0 = gp(I.w0);
    fpnl32::nl_step_row<NT, TPB, RS>(sm, row, w0, w0, gp(I.rho), RS ? gp(I.rho_s) : nullptr, 1, 1,
                                     I.np, gtables(I.nt), gp(I.slots) + 2);
  };
  if constexpr (PERS) {
    const int total = g.begin[g.n];
    for (int item = blockIdx.x; item < total; item += gridDim.x) {
      one(item);
      __syncthreads();
    }
  } else {
    one(blockIdx.x);   // one block per item: no loop state (the 4096-sample rows then need
                       // fewer registers)
  }
}

// STAGE 0..3 = G1..G4 of the mixed z-step (operands and scale slots as in the
// single-instance sequence). The stage is a template parameter: a run-time switch
// selecting the output pointer was miscompiled by hipcc (ROCm 7.2) -- its default
// case kept case 1's pointer.
template <int BM, int BN, int STAGE, int WM = 2, int WN = 2>
__global__ __launch_bounds__(WM * WN * 64, (WM * WN > 4 ? 1 : fpgemm16::TileOcc<BM, BN>::value))
void batch_gemm_kernel(const fp_batch_instance* __restrict__ insts, BatchGroup g) {
  constexpr int NT = WM * WN * 64;
  __shared__ fpgemm16::Smem<BM, BN, WM, WN> sm;
  const int total = g.begin[g.n];
  for (int it = blockIdx.x; it < total; it += gridDim.x) {
    // XCD-aware order also on a persistent grid: with gridDim.x % 8 == 0, iteration `it`
    // of any block runs on XCD it % 8, so the remap keeps each XCD on a contiguous range
    // of tiles (whole column panels) and the panel's B rows stay in that XCD's L2.
    const int item = (gridDim.x >= total || gridDim.x % 8 == 0) ? fpgemm16::xcd_remap(it, total) : it;
    const int j = group_find(g, item);
    const fp_batch_instance& I = insts[g.idx[j]];
    const int local = item - g.begin[j];
    const int tilesM = I.Nr / BM;
    //                G1        G2       G3        G4
    const float* B = gp(STAGE == 0 ? I.spec : STAGE == 2 ? I.w0 : I.w1);
    float* C = gp(STAGE == 1 ? I.w0 : STAGE == 3 ? I.spec : I.w1);
    const float alpha = (STAGE & 1) ? I.iRk : I.Rk;
    const float* rs = STAGE == 3 ? gp(I.mask_r) : nullptr;
    const float* ph = (STAGE & 1) ? nullptr : gp(I.phase);
    unsigned* slots = gp(I.slots);
    const unsigned* bin = slots + STAGE;
    unsigned* cout = STAGE == 0 ? slots + 1 : STAGE == 2 ? slots + 3 : STAGE == 3 ? slots : nullptr;
    if (STAGE == 1 && local == 0 && threadIdx.x == 0) slots[0] = 0u;
    const int tm = local % tilesM, tn = local / tilesM;
    if (tn * BN >= I.zc_lo && (tn + 1) * BN <= I.zc_hi) {
      // masked frequencies: the product is exactly zero there (phase x mask = 0 in G1 / G3,
      // zero input columns in G2 / G4); write the zeros, skip the MFMA work
      float4* Cz = reinterpret_cast<float4*>(C);
      const int ld4 = (2 * I.Nt) / 4;
      for (int i = threadIdx.x; i < BM * BN / 4; i += NT) {
        const int r = i / (BN / 4), c4 = i % (BN / 4);
        Cz[(size_t)(tm * BM + r) * ld4 + (tn * BN) / 4 + c4] = make_float4(0.f, 0.f, 0.f, 0.f);
      }
      __syncthreads();
      continue;
    }
    // phase prefetch of one 16-row block per wave where the registers allow it (the
    // 128 x 128 and 128 x 256 tiles would spill)
    constexpr int NPF = (BM == 128 && BN >= 128) ? 0 : 1;
    fpgemm16::gemm_tile<BM, BN, WM, WN, NPF>(sm, tm, tn, reinterpret_cast<const fpgemm16::h8*>(gp(I.Yh)),
                                reinterpret_cast<const fpgemm16::h8*>(gp(I.Yl)), I.y_scale_exp, B, C,
                                I.Nr, I.Nr, 2 * I.Nt, alpha, rs, ph, bin, cout);
    __syncthreads();
  }
}

// ------------------------------------------------------------------ merged passes
// One launch per pass for the whole batch: every instance's rows (or tiles) form one work
// list, whatever its Nt and STE flag (row kernels); the grid has one block per work item.
// A block finds its item's instance (uniform: the index is read with readfirstlane, so the
// descriptor loads are scalar loads), reads every descriptor field it needs before its first
// store (a later read through the descriptor pointer could alias that store and would become
// a vector load, with the pointers then held in VGPRs), and switches on the instance's Nt to
// the inlined row body of that size. All bodies share the kernel's register budget, which is
// that of the largest; no body is a separate function (a call spills its live registers to
// scratch memory around it). The heavier instances come first in the list (rebuild_groups).
template <bool HAS_STE>
struct MergedPlasmaSmem {
  static constexpr int bytes =
      cmax(cmax(PlasmaSmem<float, 512, 256, HAS_STE>::bytes, PlasmaSmem<float, 1024, 256, HAS_STE>::bytes),
           cmax(PlasmaSmem<float, 2048, 256, HAS_STE>::bytes, PlasmaSmem<float, 4096, 256, HAS_STE>::bytes));
};

template <bool HAS_STE>
__global__ __launch_bounds__(256, (PlasmaOcc<float, HAS_STE>::value)) void batch_plasma_merged(
    const fp_batch_instance* __restrict__ insts, BatchGroup g) {
  __shared__ __attribute__((aligned(16))) char smem[MergedPlasmaSmem<HAS_STE>::bytes];
  const int item = blockIdx.x;
  const int j = group_find(g, item);
  const fp_batch_instance& I = insts[__builtin_amdgcn_readfirstlane(g.idx[j])];
  const int row = item - g.begin[j];
  const int Nt = I.Nt;
  const bool ste = HAS_STE && I.enable_ste;
  const fp_plasma_params pp = I.pp;
  const float* spec = gp(I.spec);
  double *rho = gp(I.rho), *rho_s = HAS_STE ? gp(I.rho_s) : nullptr;
  const double *W = gp(I.W_lut), *Ws = gp(I.Ws_lut);
  const float* tw = gp(I.tw);
  int* status = gp(I.status);
  unsigned* slots = gp(I.slots);
  if (row == 0 && threadIdx.x < 3) slots[1 + threadIdx.x] = 0u;
#define FP_MPL(NT)                                                                             \
  if (ste)                                                                                     \
    plasma_row_body<float, NT, 256, HAS_STE, true>(smem, row, spec, rho, rho_s, nullptr, pp, W, \
                                                   Ws, tw, status);                            \
  else                                                                                         \
    plasma_row_body<float, NT, 256, false, true>(smem, row, spec, rho, nullptr, nullptr, pp, W, \
                                                 Ws, tw, status);                              \
  break;
  switch (Nt) {
    case 512: FP_MPL(512)
    case 1024: FP_MPL(1024)
    case 2048: FP_MPL(2048)
    default: FP_MPL(4096)
  }
#undef FP_MPL
}

// Nonlinear step, 512 threads: an item is 4096 / Nt rows of one instance, one per group of
// Nt / 8 threads (eight samples per thread), each group with its own LDS region. 128 VGPRs,
// two blocks (16 waves) and 160 KiB of LDS per CU.
template <int NT, bool RS>
FP_DEV void nl_rows_merged(char* smem, const fp_batch_instance& I, int grp, float* w0,
                           const double* rho, const double* rho_s, unsigned* cmax) {
  constexpr int TPB = NT / 8, RPB = 512 / TPB;
  const int sub = threadIdx.x / TPB;
  auto& sm = reinterpret_cast<fpnl32::StepSmem<NT>*>(smem)[sub];
  fpnl32::nl_step_row<NT, TPB, RS>(sm, (size_t)grp * RPB + sub, w0, w0, rho, RS ? rho_s : nullptr,
                                   1, 1, I.np, gtables(I.nt), cmax);
}

template <bool HAS_STE>
__global__ __launch_bounds__(512, 4) void batch_nl_merged(const fp_batch_instance* __restrict__ insts,
                                                          BatchGroup g) {
  __shared__ __attribute__((aligned(16))) char smem[sizeof(fpnl32::StepSmem<4096>)];
  static_assert(8 * sizeof(fpnl32::StepSmem<512>) <= sizeof(fpnl32::StepSmem<4096>), "LDS");
  const int item = blockIdx.x;
  const int j = group_find(g, item);
  const fp_batch_instance& I = insts[__builtin_amdgcn_readfirstlane(g.idx[j])];
  const int grp = item - g.begin[j];
  const int Nt = I.Nt;
  const bool ste = HAS_STE && I.enable_ste;
  float* w0 = gp(I.w0);
  const double* rho = gp(I.rho);
  const double* rho_s = HAS_STE ? gp(I.rho_s) : nullptr;
  unsigned* cmax = gp(I.slots) + 2;
#define FP_MNL(NT)                                                      \
  if (ste) nl_rows_merged<NT, HAS_STE>(smem, I, grp, w0, rho, rho_s, cmax); \
  else nl_rows_merged<NT, false>(smem, I, grp, w0, rho, rho_s, cmax);       \
  break;
  switch (Nt) {
    case 512: FP_MNL(512)
    case 1024: FP_MNL(1024)
    case 2048: FP_MNL(2048)
    default: FP_MNL(4096)
  }
#undef FP_MNL
}

// GEMM stage, 512 threads, one 256 x 128 output tile (4 x 2 waves) per block, for instances
// with Nr % 256 == 0 (fp_batch_set_merged checks it); same operands, scale slots and
// masked-tile skipping as batch_gemm_kernel. XCD-aware item order (whole column panels of an
// instance per XCD, so a panel's B rows stay in that XCD's L2).
template <int STAGE>
__global__ __launch_bounds__(512, 1) void batch_gemm_merged(const fp_batch_instance* __restrict__ insts,
                                                            BatchGroup g) {
  constexpr int BM = 256, BN = 128;
  __shared__ fpgemm16::Smem<BM, BN, 4, 2> sm;
  const int total = g.begin[g.n];
  const int item = fpgemm16::xcd_remap(blockIdx.x, total);
  const int j = group_find(g, item);
  const fp_batch_instance& I = insts[__builtin_amdgcn_readfirstlane(g.idx[j])];
  const int local = item - g.begin[j];
  const int Nr = I.Nr, Nt = I.Nt, zc_lo = I.zc_lo, zc_hi = I.zc_hi, yse = I.y_scale_exp;
  const auto* Yh = reinterpret_cast<const fpgemm16::h8*>(gp(I.Yh));
  const auto* Yl = reinterpret_cast<const fpgemm16::h8*>(gp(I.Yl));
  //                G1        G2       G3        G4
  const float* B = gp(STAGE == 0 ? I.spec : STAGE == 2 ? I.w0 : I.w1);
  float* C = gp(STAGE == 1 ? I.w0 : STAGE == 3 ? I.spec : I.w1);
  const float alpha = (STAGE & 1) ? I.iRk : I.Rk;
  const float* rs = STAGE == 3 ? gp(I.mask_r) : nullptr;
  const float* ph = (STAGE & 1) ? nullptr : gp(I.phase);
  unsigned* slots = gp(I.slots);
  const unsigned* bin = slots + STAGE;
  unsigned* cout = STAGE == 0 ? slots + 1 : STAGE == 2 ? slots + 3 : STAGE == 3 ? slots : nullptr;
  if (STAGE == 1 && local == 0 && threadIdx.x == 0) slots[0] = 0u;
  const int tilesM = Nr / BM;
  const int tm = local % tilesM, tn = local / tilesM;
  if (tn * BN >= zc_lo && (tn + 1) * BN <= zc_hi) {
    float4* Cz = reinterpret_cast<float4*>(C);
    const int ld4 = (2 * Nt) / 4;
    for (int i = threadIdx.x; i < BM * BN / 4; i += 512) {
      const int r = i / (BN / 4), c4 = i % (BN / 4);
      Cz[(size_t)(tm * BM + r) * ld4 + (tn * BN) / 4 + c4] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    return;
  }
  fpgemm16::gemm_tile<BM, BN, 4, 2, 1>(sm, tm, tn, Yh, Yl, yse, B, C, Nr, Nr, 2 * Nt, alpha, rs,
                                      ph, bin, cout);
}

}  // namespace

// Host-side plan of the batched form: descriptors (host copy for validation and
// grouping, device copy read by the kernels), per-instance launch configuration
// and the launch groups derived from it.
struct fp_batch_plan {
  int n = 0;
  int device = 0;
  std::vector<fp_batch_instance> host;
  fp_batch_instance* dev = nullptr;
  std::vector<int> bm, bn, nl_tpb;
  int grid[3] = {0, 0, 0};
  struct Launch {
    int Nt, ste, a, b;          // row kernels: a = threads per block; GEMM: a x b tile
    BatchGroup g;
  };
  std::vector<Launch> plasma, nl, gemm;           // grouped launches
  std::vector<Launch> plasma_m, nl_m, gemm_m;     // merged work lists (one launch per pass)
  bool dirty = true;
  unsigned merged = 0;        // bit FP_PASS_x: that pass is one launch for the whole batch
};

namespace {

void push_group(std::vector<fp_batch_plan::Launch>& out, const fp_batch_plan::Launch& key,
                const std::vector<int>& members, const std::vector<int>& items) {
  for (size_t s = 0; s < members.size(); s += kGroupMax) {
    fp_batch_plan::Launch L = key;
    L.g.n = 0;
    L.g.begin[0] = 0;
    for (size_t k = s; k < members.size() && k < s + kGroupMax; ++k) {
      L.g.idx[L.g.n] = members[k];
      L.g.begin[L.g.n + 1] = L.g.begin[L.g.n] + items[k];
      ++L.g.n;
    }
    out.push_back(L);
  }
}

// Groups: plasma by (Nt, STE); nonlinear by (Nt, STE, threads); GEMM by tile shape.
// Instances keep their order inside a group.
void rebuild_groups(fp_batch_plan& P) {
  P.plasma.clear();
  P.nl.clear();
  P.gemm.clear();
  P.plasma_m.clear();
  P.nl_m.clear();
  P.gemm_m.clear();
  if (P.merged) {
    // every instance in one group per pass, heavier instances first
    std::vector<int> order(P.n);
    for (int i = 0; i < P.n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
      const fp_batch_instance& a = P.host[x];
      const fp_batch_instance& b = P.host[y];
      return a.Nt != b.Nt ? a.Nt > b.Nt : a.Nr > b.Nr;
    });
    bool any_ste = false;
    for (int i = 0; i < P.n; ++i) any_ste = any_ste || P.host[i].enable_ste;
    std::vector<int> rows, groups, tiles;
    for (int i : order) {
      const fp_batch_instance& I = P.host[i];
      rows.push_back(I.Nr);
      groups.push_back(I.Nr / (4096 / I.Nt));
      tiles.push_back((I.Nr / 256) * (2 * I.Nt / 128));   // 256 x 128 tiles
    }
    fp_batch_plan::Launch key{};
    key.ste = any_ste;
    push_group(P.plasma_m, key, order, rows);
    push_group(P.nl_m, key, order, groups);
    push_group(P.gemm_m, key, order, tiles);
  }
  std::vector<std::array<int, 4>> keys;
  for (int kind = 0; kind < 3; ++kind) {
    keys.clear();
    for (int i = 0; i < P.n; ++i) {
      const fp_batch_instance& I = P.host[i];
      std::array<int, 4> k = kind == 0   ? std::array<int, 4>{I.Nt, I.enable_ste, 256, 0}
                             : kind == 1 ? std::array<int, 4>{I.Nt, I.enable_ste, P.nl_tpb[i], 0}
                                         : std::array<int, 4>{0, 0, P.bm[i], P.bn[i]};
      if (std::find(keys.begin(), keys.end(), k) == keys.end()) keys.push_back(k);
    }
    // larger work first (GEMM: taller tiles first) so persistent tails are short
    std::stable_sort(keys.begin(), keys.end(), [](const std::array<int, 4>& x,
                                                  const std::array<int, 4>& y) {
      return x[0] * 1000 + x[2] * 4 + x[3] > y[0] * 1000 + y[2] * 4 + y[3];
    });
    for (const auto& k : keys) {
      std::vector<int> members, items;
      for (int i = 0; i < P.n; ++i) {
        const fp_batch_instance& I = P.host[i];
        const bool match = kind == 0   ? (I.Nt == k[0] && I.enable_ste == k[1])
                           : kind == 1 ? (I.Nt == k[0] && I.enable_ste == k[1] && P.nl_tpb[i] == k[2])
                                       : (P.bm[i] == k[2] && P.bn[i] == k[3]);
        if (!match) continue;
        members.push_back(i);
        items.push_back(kind < 2 ? I.Nr : (I.Nr / k[2]) * (2 * I.Nt / k[3]));
      }
      fp_batch_plan::Launch key{};
      key.Nt = k[0];
      key.ste = k[1];
      key.a = k[2];
      key.b = k[3];
      push_group(kind == 0 ? P.plasma : kind == 1 ? P.nl : P.gemm, key, members, items);
    }
  }
  P.dirty = false;
}

inline dim3 grid_of(const BatchGroup& g, int persistent) {
  const int total = g.begin[g.n];
  return dim3(persistent > 0 && persistent < total ? persistent : total);
}

inline bool persistent(const BatchGroup& g, int blocks) { return blocks > 0 && blocks < g.begin[g.n]; }

template <int NT, bool STE>
void launch_batch_plasma_t(const fp_batch_plan& P, const fp_batch_plan::Launch& L, hipStream_t st) {
  const dim3 grid = grid_of(L.g, P.grid[0]);
  if (persistent(L.g, P.grid[0]))
    hipLaunchKernelGGL((batch_plasma_kernel<NT, 256, STE, true>), grid, dim3(256), 0, st, P.dev, L.g);
  else
    hipLaunchKernelGGL((batch_plasma_kernel<NT, 256, STE, false>), grid, dim3(256), 0, st, P.dev, L.g);
}

int launch_batch_plasma(const fp_batch_plan& P, const fp_batch_plan::Launch& L, hipStream_t st) {
#define FP_BP(NT)                                                                               \
  if (L.ste) launch_batch_plasma_t<NT, true>(P, L, st);                                         \
  else launch_batch_plasma_t<NT, false>(P, L, st);                                              \
  break;
  switch (L.Nt) {
    case 512: FP_BP(512)
    case 1024: FP_BP(1024)
    case 2048: FP_BP(2048)
    case 4096: FP_BP(4096)
    default: return fail(FP_E_SHAPE, "batch plasma: bad Nt");
  }
#undef FP_BP
  return check_launch("batch_plasma");
}

template <int NT, int TPB>
void launch_batch_nl_t(const fp_batch_plan& P, const fp_batch_plan::Launch& L, hipStream_t st) {
  const dim3 grid = grid_of(L.g, P.grid[2]);
  const bool pers = persistent(L.g, P.grid[2]);
  if (L.ste) {
    if (pers) hipLaunchKernelGGL((batch_nl_kernel<NT, TPB, true, true>), grid, dim3(TPB), 0, st, P.dev, L.g);
    else hipLaunchKernelGGL((batch_nl_kernel<NT, TPB, true, false>), grid, dim3(TPB), 0, st, P.dev, L.g);
  } else {
    if (pers) hipLaunchKernelGGL((batch_nl_kernel<NT, TPB, false, true>), grid, dim3(TPB), 0, st, P.dev, L.g);
    else hipLaunchKernelGGL((batch_nl_kernel<NT, TPB, false, false>), grid, dim3(TPB), 0, st, P.dev, L.g);
  }
}

int launch_batch_nl(const fp_batch_plan& P, const fp_batch_plan::Launch& L, hipStream_t st) {
  const bool wide = L.a == L.Nt / 4;
#define FP_BN(NT)                                                              \
  if (wide) launch_batch_nl_t<NT, NT / 4>(P, L, st);                           \
  else launch_batch_nl_t<NT, NT / 8>(P, L, st);                                \
  break;
  switch (L.Nt) {
    case 512: FP_BN(512)
    case 1024: FP_BN(1024)
    case 2048: FP_BN(2048)
    case 4096:
      if (L.a == 256) launch_batch_nl_t<4096, 256>(P, L, st);
      else if (wide) launch_batch_nl_t<4096, 1024>(P, L, st);
      else launch_batch_nl_t<4096, 512>(P, L, st);
      break;
    default: return fail(FP_E_SHAPE, "batch nonlinear: bad Nt");
  }
#undef FP_BN
  return check_launch("batch_nl_step");
}

template <int BM, int BN, int WM = 2, int WN = 2>
void launch_batch_gemm_t(const fp_batch_plan& P, const fp_batch_plan::Launch& L, int stage,
                         hipStream_t st) {
  const dim3 grid = grid_of(L.g, P.grid[1]);
  const dim3 blk(WM * WN * 64);
  switch (stage) {
    case 0: hipLaunchKernelGGL((batch_gemm_kernel<BM, BN, 0, WM, WN>), grid, blk, 0, st, P.dev, L.g); break;
    case 1: hipLaunchKernelGGL((batch_gemm_kernel<BM, BN, 1, WM, WN>), grid, blk, 0, st, P.dev, L.g); break;
    case 2: hipLaunchKernelGGL((batch_gemm_kernel<BM, BN, 2, WM, WN>), grid, blk, 0, st, P.dev, L.g); break;
    default: hipLaunchKernelGGL((batch_gemm_kernel<BM, BN, 3, WM, WN>), grid, blk, 0, st, P.dev, L.g); break;
  }
}

int launch_batch_gemm(const fp_batch_plan& P, const fp_batch_plan::Launch& L, int stage,
                      hipStream_t st) {
  if (stage < 0 || stage > 3) return fail(FP_E_ARG, "batch GEMM: bad stage");
  if (L.a == 256 && L.b == 128) launch_batch_gemm_t<256, 128, 4, 2>(P, L, stage, st);
  else if (L.a == 128 && L.b == 256) launch_batch_gemm_t<128, 256>(P, L, stage, st);
  else if (L.a == 128 && L.b == 128) launch_batch_gemm_t<128, 128>(P, L, stage, st);
  else if (L.a == 128 && L.b == 64) launch_batch_gemm_t<128, 64>(P, L, stage, st);
  else if (L.a == 64 && L.b == 128) launch_batch_gemm_t<64, 128>(P, L, stage, st);
  else if (L.a == 64 && L.b == 64) launch_batch_gemm_t<64, 64>(P, L, stage, st);
  else return fail(FP_E_ARG, "batch GEMM: bad tile");
  return check_launch("batch_hankel_gemm");
}

int run_pass_merged(fp_batch_plan& P, int pass, hipStream_t st) {
  // one block per work item
  auto grid = [&](const BatchGroup& g) {
    const int total = g.begin[g.n];
    return dim3(total < 1 ? 1 : total);
  };
  switch (pass) {
    case FP_PASS_PLASMA:
      for (const auto& L : P.plasma_m) {
        if (L.ste)
          hipLaunchKernelGGL((batch_plasma_merged<true>), grid(L.g), dim3(256), 0, st, P.dev, L.g);
        else
          hipLaunchKernelGGL((batch_plasma_merged<false>), grid(L.g), dim3(256), 0, st, P.dev,
                             L.g);
      }
      return check_launch("batch_plasma_merged");
    case FP_PASS_NL:
      for (const auto& L : P.nl_m) {
        if (L.ste)
          hipLaunchKernelGGL((batch_nl_merged<true>), grid(L.g), dim3(512), 0, st, P.dev, L.g);
        else
          hipLaunchKernelGGL((batch_nl_merged<false>), grid(L.g), dim3(512), 0, st, P.dev, L.g);
      }
      return check_launch("batch_nl_merged");
    case FP_PASS_G1: case FP_PASS_G2: case FP_PASS_G3: case FP_PASS_G4: {
      for (const auto& L : P.gemm_m) {
        const dim3 gr = grid(L.g);
        switch (pass) {
          case FP_PASS_G1: hipLaunchKernelGGL((batch_gemm_merged<0>), gr, dim3(512), 0, st, P.dev, L.g); break;
          case FP_PASS_G2: hipLaunchKernelGGL((batch_gemm_merged<1>), gr, dim3(512), 0, st, P.dev, L.g); break;
          case FP_PASS_G3: hipLaunchKernelGGL((batch_gemm_merged<2>), gr, dim3(512), 0, st, P.dev, L.g); break;
          default: hipLaunchKernelGGL((batch_gemm_merged<3>), gr, dim3(512), 0, st, P.dev, L.g); break;
        }
      }
      return check_launch("batch_gemm_merged");
    }
    default: return fail(FP_E_ARG, "fp_batch_pass: unknown pass");
  }
}

int run_pass(fp_batch_plan& P, int pass, hipStream_t st) {
  if (P.dirty) rebuild_groups(P);
  if (pass >= 0 && pass < 6 && ((P.merged >> pass) & 1u)) return run_pass_merged(P, pass, st);
  int rc = FP_OK;
  switch (pass) {
    case FP_PASS_PLASMA:
      for (const auto& L : P.plasma) if ((rc = launch_batch_plasma(P, L, st))) return rc;
      return FP_OK;
    case FP_PASS_NL:
      for (const auto& L : P.nl) if ((rc = launch_batch_nl(P, L, st))) return rc;
      return FP_OK;
    case FP_PASS_G1: case FP_PASS_G2: case FP_PASS_G3: case FP_PASS_G4: {
      const int stage = pass == FP_PASS_G1 ? 0 : pass == FP_PASS_G2 ? 1 : pass == FP_PASS_G3 ? 2 : 3;
      for (const auto& L : P.gemm) if ((rc = launch_batch_gemm(P, L, stage, st))) return rc;
      return FP_OK;
    }
    default: return fail(FP_E_ARG, "fp_batch_pass: unknown pass");
  }
}

bool valid_nt(int Nt) { return Nt == 512 || Nt == 1024 || Nt == 2048 || Nt == 4096; }

// Eight samples per thread. For Nt <= 2048 that is fft_rows_kernel's block size, so the
// time-domain row is bitwise the one fp_fft_rows returns; at Nt = 4096 (512 threads against
// fft_rows_kernel's 256) it differs only by transform rounding, and the kernel stays at 126
// VGPRs (4 waves per SIMD): a record costs 52 instead of 59 us (mixed) at 1024 x 4096.
template <typename T, int NT, int TPB>
int launch_record_history(const void* spec, const void* tw, const double* rho, const double* rho_s,
                          const fp_nl_params& p, const fp_nl_tables& t, double dt, int nbins,
                          const int* bin_lo, const int* bin_hi, const fp_record_dst& d,
                          hipStream_t st) {
  hipLaunchKernelGGL((fprec::record_history_kernel<T, NT, TPB>), dim3(p.Nr), dim3(TPB), 0, st,
                     static_cast<const T*>(spec), static_cast<const T*>(tw), rho, rho_s, p, t, dt,
                     nbins, bin_lo, bin_hi, d);
  return check_launch("record_history");
}

template <typename T>
int dispatch_record_history(const void* spec, const void* tw, const double* rho,
                            const double* rho_s, const fp_nl_params& p, const fp_nl_tables& t,
                            double dt, int nbins, const int* bin_lo, const int* bin_hi,
                            const fp_record_dst& d, hipStream_t st) {
  switch (p.Nt) {
    case 512: return launch_record_history<T, 512, 64>(spec, tw, rho, rho_s, p, t, dt, nbins, bin_lo, bin_hi, d, st);
    case 1024: return launch_record_history<T, 1024, 128>(spec, tw, rho, rho_s, p, t, dt, nbins, bin_lo, bin_hi, d, st);
    case 2048: return launch_record_history<T, 2048, 256>(spec, tw, rho, rho_s, p, t, dt, nbins, bin_lo, bin_hi, d, st);
    case 4096: return launch_record_history<T, 4096, 512>(spec, tw, rho, rho_s, p, t, dt, nbins, bin_lo, bin_hi, d, st);
    default: return fail(FP_E_SHAPE, "fp_record_history: Nt must be 512, 1024, 2048 or 4096");
  }
}

int record_history(const void* spec, int spec_c64, const void* tw, const double* rho,
                   const double* rho_s, const fp_nl_params* p, const fp_nl_tables* t, double dt,
                   int nbins, const int* bin_lo, const int* bin_hi, const fp_record_dst& d,
                   hipStream_t st);
int record_check(const void* spec, const void* tw, const double* rho, const fp_nl_params* p,
                 const fp_nl_tables* t, int nbins, const int* bin_lo, const int* bin_hi,
                 const fp_record_dst& dd, bool need_partial);

// Scratch of fp_record_history_noalloc: the 3 x Nr FP64 per-row partials of one record
// (3 x 16384 x 8 B = 384 KiB per device, held until the process ends). In host code
// `scratch` is only a stand-in symbol: its device address comes from hipGetSymbolAddress,
// once per device, and only that address is passed to kernels.
static __device__ __align__(256) double scratch[3 * FP_NOALLOC_MAX_NR];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up on the first call there).
int record_scratch(double** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    g_err = std::string("fp_record_history_noalloc: hipGetDevice: ") + hipGetErrorString(e);
    return (int)e;
  }
  if (dev < 0 || dev >= kMaxDevices) return fail(FP_E_ARG, "fp_record_history_noalloc: device index");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* a = nullptr;
    e = hipGetSymbolAddress(&a, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !a) {
      g_err = std::string("fp_record_history_noalloc: hipGetSymbolAddress: ") +
              hipGetErrorString(e);
      return e != hipSuccess ? (int)e : FP_E_ARG;
    }
    g_scratch_dev[dev] = static_cast<double*>(a);
  }
  *out = g_scratch_dev[dev];
  return FP_OK;
}

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
  if (p->Nr <= 0) return fail(FP_E_SHAPE, "fp_nl_split: Nr must be positive");
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
  if (p->Nr <= 0) return fail(FP_E_SHAPE, "fp_nl_step: Nr must be positive");
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
  if (p->Nr <= 0) return fail(FP_E_SHAPE, "fp_record_rows: Nr must be positive");
  if (nbins > 0 && (!bin_lo || !bin_hi || !bins)) return fail(FP_E_ARG, "fp_record_rows: bins");
  hipLaunchKernelGGL((fpnl::record_kernel<256>), dim3(p->Nr), dim3(256), 0, S(stream), u, rho,
                     rho_s, rows_out, *p, *t, dt, nbins, bin_lo, bin_hi, bins);
  return check_launch("record_rows");
}

int fp_record_history(const void* spec, int spec_c64, const void* tw, const double* rho,
                      const double* rho_s, const fp_nl_params* p, const fp_nl_tables* t,
                      double dt, int nbins, const int* bin_lo, const int* bin_hi,
                      const fp_record_dst* d, void* stream) {
  if (!d) return fail(FP_E_ARG, "fp_record_history: null pointer");
  if (!d->partial) return fail(FP_E_ARG, "fp_record_history: null destination");
  return record_history(spec, spec_c64, tw, rho, rho_s, p, t, dt, nbins, bin_lo, bin_hi, *d,
                        S(stream));
}

int fp_record_history_noalloc(const void* spec, int spec_c64, const void* tw, const double* rho,
                              const double* rho_s, const fp_nl_params* p, const fp_nl_tables* t,
                              double dt, int nbins, const int* bin_lo, const int* bin_hi,
                              const fp_record_dst* d, void* stream) {
  if (!d || !p) return fail(FP_E_ARG, "fp_record_history_noalloc: null pointer");
  if (p->Nr > FP_NOALLOC_MAX_NR)
    return fail(FP_E_SHAPE, "fp_record_history_noalloc: Nr exceeds FP_NOALLOC_MAX_NR (16384)");
  int rc = record_check(spec, tw, rho, p, t, nbins, bin_lo, bin_hi, *d, false);
  if (rc != FP_OK) return rc;
  double* partial = nullptr;
  if ((rc = record_scratch(&partial)) != FP_OK) return rc;
  fp_record_dst dl = *d;
  dl.partial = partial;
  return record_history(spec, spec_c64, tw, rho, rho_s, p, t, dt, nbins, bin_lo, bin_hi, dl,
                        S(stream));
}

}  // extern "C"

namespace {

// Host-side validation of a record's arguments (no device call).
int record_check(const void* spec, const void* tw, const double* rho, const fp_nl_params* p,
                 const fp_nl_tables* t, int nbins, const int* bin_lo, const int* bin_hi,
                 const fp_record_dst& dd, bool need_partial) {
  const fp_record_dst* d = &dd;
  if (!spec || !tw || !rho || !p || !t)
    return fail(FP_E_ARG, "fp_record_history: null pointer");
  if (p->Nr <= 0) return fail(FP_E_SHAPE, "fp_record_history: Nr must be positive");
  if (!d->fluence || !d->rho || !d->rho_s || !d->ipeak || !d->absorbed || !d->rho_onaxis ||
      !d->rho_s_onaxis || !d->I_onaxis || !d->r_w || (need_partial && !d->partial) ||
      !d->imax || !d->dedz_photo || !d->dedz_plasma)
    return fail(FP_E_ARG, "fp_record_history: null destination");
  if (nbins < 0 || (nbins > 0 && (!bin_lo || !bin_hi || !d->bins)))
    return fail(FP_E_ARG, "fp_record_history: bins");
  if (d->rho_at && (d->it_snap < 0 || d->it_snap >= p->Nt))
    return fail(FP_E_ARG, "fp_record_history: it_snap outside [0, Nt)");
  if (d->t_stride < 0) return fail(FP_E_ARG, "fp_record_history: t_stride < 0");
  if (d->t_stride > 0) {
    if (d->r_stride <= 0) return fail(FP_E_ARG, "fp_record_history: r_stride must be positive");
    if (d->Nt_sub != (p->Nt - 1) / d->t_stride + 1)
      return fail(FP_E_ARG, "fp_record_history: Nt_sub != (Nt-1)/t_stride + 1");
    if (!d->rho_cube || !d->rho_s_cube || !d->I_cube)
      return fail(FP_E_ARG, "fp_record_history: null cube destination");
  }
  if (p->Nt != 512 && p->Nt != 1024 && p->Nt != 2048 && p->Nt != 4096)
    return fail(FP_E_SHAPE, "fp_record_history: Nt must be 512, 1024, 2048 or 4096");
  return FP_OK;
}

int record_history(const void* spec, int spec_c64, const void* tw, const double* rho,
                   const double* rho_s, const fp_nl_params* p, const fp_nl_tables* t, double dt,
                   int nbins, const int* bin_lo, const int* bin_hi, const fp_record_dst& dd,
                   hipStream_t st) {
  const fp_record_dst* d = &dd;
  int rc0 = record_check(spec, tw, rho, p, t, nbins, bin_lo, bin_hi, dd, true);
  if (rc0 != FP_OK) return rc0;
  const int rc = spec_c64 ? dispatch_record_history<float>(spec, tw, rho, rho_s, *p, *t, dt, nbins,
                                                           bin_lo, bin_hi, *d, st)
                          : dispatch_record_history<double>(spec, tw, rho, rho_s, *p, *t, dt,
                                                            nbins, bin_lo, bin_hi, *d, st);
  if (rc != FP_OK) return rc;
  hipLaunchKernelGGL(fprec::record_finish_kernel, dim3(1), dim3(fprec::FIN_TPB), 0, st, p->Nr, dt,
                     *d);
  return check_launch("record_finish");
}

}  // namespace

extern "C" {

// ------------------------------------------------------------------ mixed precision
// The FP32 PCHIP table is read one float4 per interval.
static bool pcoef_misaligned(const fp_nl_tables_f32* t) {
  return reinterpret_cast<uintptr_t>(t->pcoef_s) % 16 != 0;
}

int fp_absmax_f32(const float* x, long long n, unsigned* slot, void* stream) {
  if (!x || !slot || n < 0) return fail(FP_E_ARG, "fp_absmax_f32: bad argument");
  if (reinterpret_cast<uintptr_t>(x) % 16) return fail(FP_E_ARG, "fp_absmax_f32: x must be 16-byte aligned");
  long long blocks = (n / 4 + 255) / 256;
  if (blocks > 512) blocks = 512;
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

}  // extern "C"

namespace {

// Compute units of the current device (cached per device; 256 if the query fails).
int device_cus() {
  static int cus[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (!cus[dev]) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        v <= 0)
      v = 256;
    cus[dev] = v;
  }
  return cus[dev];
}

}  // namespace

extern "C" {

int fp_hankel_gemm_mixed(const void* Yh, const void* Yl, int y_scale_exp, const float* B,
                         float* C, int Nr, int Nt, float alpha, const float* rowscale,
                         const float* phase, const unsigned* bmax_in, unsigned* cmax_out,
                         unsigned* reset_slot, void* stream) {
  if (!Yh || !Yl || !B || !C || !bmax_in) return fail(FP_E_ARG, "fp_hankel_gemm_mixed: null pointer");
  if (B == C) return fail(FP_E_ARG, "fp_hankel_gemm_mixed: C must not alias B");
  const int Ncols = 2 * Nt;
  if (Nr <= 0 || Nr % 128 || Ncols % 128)
    return fail(FP_E_SHAPE, "fp_hankel_gemm_mixed: needs Nr % 128 == 0 and 2 Nt % 128 == 0");
  // Tile by the number of tiles against the CUs (every tile shape gives each output the same
  // K order, so the results are bitwise equal). 256 x 128 tiles of 8 waves, one block per CU,
  // when they fill the CUs: 5 % faster than 128 x 128 tiles at 1024 x 4096 (each B panel is
  // staged by half as many blocks). Smaller problems take tiles of 4 waves, two blocks per CU:
  // 128 x 128 when they fill both slots of every CU, else 128 x 64 when they fill the CUs,
  // else 64 x 64 (mixed step 512 x 2048: 0.194 -> 0.158 ms, 256 x 1024: 0.120 -> 0.090 ms,
  // 128 x 512: 0.088 -> 0.069 ms against 256 x 128 / 128 x 128 tiles).
  const auto* yh = reinterpret_cast<const fpgemm16::h8*>(Yh);
  const auto* yl = reinterpret_cast<const fpgemm16::h8*>(Yl);
  constexpr int NPF = FP_MIXED_GEMM_NPF;
  const int cus = device_cus();
  auto tiles = [&](int bm, int bn) { return (Nr / bm) * (Ncols / bn); };
#define FP_MIXED_TILE(BM, BN, WM, WN)                                                          \
  hipLaunchKernelGGL((fpgemm16::hankel_gemm_f16x3<BM, BN, WM, WN, NPF>), dim3(tiles(BM, BN)),  \
                     dim3(WM * WN * 64), 0, S(stream), yh, yl, y_scale_exp, B, C, Nr, Nr, Ncols, \
                     alpha, rowscale, phase, bmax_in, cmax_out, reset_slot)
  if (Nr % 256 == 0 && tiles(256, 128) >= cus) FP_MIXED_TILE(256, 128, 4, 2);
  else if (tiles(128, 128) >= 2 * cus) FP_MIXED_TILE(128, 128, 2, 2);
  else if (tiles(128, 64) >= cus) FP_MIXED_TILE(128, 64, 2, 2);
  else FP_MIXED_TILE(64, 64, 2, 2);
#undef FP_MIXED_TILE
  return check_launch("hankel_gemm_mixed");
}

}  // extern "C"

namespace {

// Column band of the single-instance Hankel GEMMs: [zc_lo, zc_hi) are real columns whose
// output is known to be zero (the masked frequencies). The band keeps the 16-column blocks
// outside the largest 16-aligned part of that range; its blocks are split in tiles of TBK
// blocks so that the tiles fill the CUs once (TBK = 4 .. 7; 8, i.e. the full tile width,
// gains nothing and returns false). Padding blocks needed to fill the last tile are taken
// from the masked range (they compute zeros).
struct BandPlan {
  int tbk, slots;
  fpgemm16::ColBand band;
  int zlo, zhi;   // real columns left to zero
};

// tile_rows x (TBK 16-column blocks) tiles, `per_cu` blocks resident per CU, TBK in
// [min_tbk, 7]. With whole_width (and no usable zero range) the tiles cover all columns:
// TBK is then lowered to a divisor of the column blocks, since there is no masked range to
// take padding blocks from.
bool plan_band(int Nr, int Ncols, int zc_lo, int zc_hi, int cus, BandPlan* bp,
               int tile_rows = 256, int per_cu = 1, int min_tbk = 4, bool whole_width = false) {
  if (Nr % tile_rows || Ncols % 16) return false;
  int a16 = Ncols, b16 = Ncols;
  if (zc_lo >= 0 && zc_hi <= Ncols && zc_hi > zc_lo) {
    a16 = (zc_lo + 15) / 16 * 16;
    b16 = zc_hi / 16 * 16;
  }
  const bool banded = b16 - a16 >= 16;
  if (!banded && !whole_width) return false;
  if (!banded) a16 = b16 = Ncols;
  const int Lb = (a16 + (Ncols - b16)) / 16;
  const int per_panel = cus * per_cu / (Nr / tile_rows);
  if (per_panel < 1 || Lb < 1) return false;
  int tbk = (Lb + per_panel - 1) / per_panel;
  if (tbk < min_tbk || tbk > 7) return false;
  if (!banded)
    while (Lb % tbk) --tbk;
  const int slots = (Lb + tbk - 1) / tbk;
  const int va = a16 + (slots * tbk - Lb) * 16;
  if (va > b16) return false;
  bp->tbk = tbk;
  bp->slots = slots;
  bp->band = fpgemm16::ColBand{va, b16 - va};
  bp->zlo = va;
  bp->zhi = b16;
  return true;
}

}  // namespace

extern "C" {

int fp_hankel_gemm_band(const double* Yt, const double* B, double* C, int Nr, int Nt,
                        double alpha, const double* rowscale, const double* phase, int zc_lo,
                        int zc_hi, int zero_fill, void* stream) {
  if (!Yt || !B || !C || Nr <= 0 || Nt <= 0) return fail(FP_E_ARG, "fp_hankel_gemm_band: bad argument");
  if (B == C) return fail(FP_E_ARG, "fp_hankel_gemm_band: C must not alias B");
  const int Ncols = 2 * Nt;
  if (zc_lo < 0 || zc_hi > Ncols || (zc_hi > 0 && zc_hi < zc_lo))
    return fail(FP_E_ARG, "fp_hankel_gemm_band: zero-column range outside [0, 2 Nt)");
  BandPlan bp;
  // 128-row tiles of TBK = 1 .. 7 blocks, also over the whole width when there is no zero
  // range: the band kernel's per-element arithmetic is hankel_gemm_f64's, and narrow tiles
  // fill the CUs at sizes where 128 x 128 tiles do not (512 x 2048: 128 tiles for 512 block
  // slots)
  if (!plan_band(Nr, Ncols, zc_lo, zc_hi, device_cus(), &bp, fpgemm::BM, 2, 1, true))
    return fp_hankel_gemm(Yt, B, C, Nr, Nt, alpha, rowscale, phase, stream);
  const int zlo = zero_fill ? bp.zlo : 0, zhi = zero_fill ? bp.zhi : 0;
  const dim3 grid((Nr / fpgemm::BM) * bp.slots);
  const fpgemm::Band64 band{bp.band.va, bp.band.shift};
#define FP_BAND64(TB)                                                                          \
  hipLaunchKernelGGL((fpgemm::hankel_gemm_f64_band<TB>), grid, dim3(fpgemm::NTHR), 0, S(stream), \
                     Yt, B, C, Nr, Nr, Ncols, alpha, rowscale, phase, bp.slots, band, zlo, zhi); \
  break;
  switch (bp.tbk) {
    case 1: FP_BAND64(1)
    case 2: FP_BAND64(2)
    case 3: FP_BAND64(3)
    case 4: FP_BAND64(4)
    case 5: FP_BAND64(5)
    case 6: FP_BAND64(6)
    default: FP_BAND64(7)
  }
#undef FP_BAND64
  return check_launch("hankel_gemm_band");
}

int fp_hankel_gemm_mixed_band(const void* Yh, const void* Yl, int y_scale_exp, const float* B,
                              float* C, int Nr, int Nt, float alpha, const float* rowscale,
                              const float* phase, const unsigned* bmax_in, unsigned* cmax_out,
                              unsigned* reset_slot, int zc_lo, int zc_hi, int zero_fill,
                              void* stream) {
  if (!Yh || !Yl || !B || !C || !bmax_in)
    return fail(FP_E_ARG, "fp_hankel_gemm_mixed_band: null pointer");
  if (B == C) return fail(FP_E_ARG, "fp_hankel_gemm_mixed_band: C must not alias B");
  const int Ncols = 2 * Nt;
  if (Nr <= 0 || Nr % 128 || Ncols % 128)
    return fail(FP_E_SHAPE, "fp_hankel_gemm_mixed_band: needs Nr % 128 == 0 and 2 Nt % 128 == 0");
  if (zc_lo < 0 || zc_hi > Ncols || (zc_hi > 0 && zc_hi < zc_lo))
    return fail(FP_E_ARG, "fp_hankel_gemm_mixed_band: zero-column range outside [0, 2 Nt)");
  BandPlan bp;
  if (!plan_band(Nr, Ncols, zc_lo, zc_hi, device_cus(), &bp))
    return fp_hankel_gemm_mixed(Yh, Yl, y_scale_exp, B, C, Nr, Nt, alpha, rowscale, phase,
                                bmax_in, cmax_out, reset_slot, stream);
  const auto* yh = reinterpret_cast<const fpgemm16::h8*>(Yh);
  const auto* yl = reinterpret_cast<const fpgemm16::h8*>(Yl);
  const int zlo = zero_fill ? bp.zlo : 0, zhi = zero_fill ? bp.zhi : 0;
  const dim3 grid((Nr / 256) * bp.slots);
  if (!phase) {
    // no phase (G2 / G4): the K-split tile (A fragments from global memory, two partial sums)
#define FP_KSPLIT(TB)                                                                          \
  hipLaunchKernelGGL((fpgemm16::hankel_gemm_f16x3_ksplit<TB>), grid, dim3(512), 0, S(stream),   \
                     yh, yl, y_scale_exp, B, C, Nr, Nr, Ncols, alpha, rowscale, bmax_in,          \
                     cmax_out, reset_slot, bp.slots, bp.band, zlo, zhi);                          \
  break;
    switch (bp.tbk) {
      case 4: FP_KSPLIT(4)
      case 5: FP_KSPLIT(5)
      case 6: FP_KSPLIT(6)
      default: FP_KSPLIT(7)
    }
#undef FP_KSPLIT
    return check_launch("hankel_gemm_mixed_band (K split)");
  }
  constexpr int NPF = FP_MIXED_GEMM_NPF;
#define FP_BAND(TB)                                                                            \
  hipLaunchKernelGGL((fpgemm16::hankel_gemm_f16x3_band<TB, NPF>), grid, dim3(512), 0, S(stream), \
                     yh, yl, y_scale_exp, B, C, Nr, Nr, Ncols, alpha, rowscale, phase, bmax_in,  \
                     cmax_out, reset_slot, bp.slots, bp.band, zlo, zhi);                         \
  break;
  switch (bp.tbk) {
    case 4: FP_BAND(4)
    case 5: FP_BAND(5)
    case 6: FP_BAND(6)
    default: FP_BAND(7)
  }
#undef FP_BAND
  return check_launch("hankel_gemm_mixed_band");
}

int fp_nl_split_c64(const float* u, const double* rho, const double* rho_s, float* rhs,
                    float* alpha, const fp_nl_params* p, const fp_nl_tables_f32* t,
                    void* stream) {
  if (!u || !rho || !rhs || !p || !t) return fail(FP_E_ARG, "fp_nl_split_c64: null pointer");
  if (p->Nr <= 0) return fail(FP_E_SHAPE, "fp_nl_split_c64: Nr must be positive");
  if (pcoef_misaligned(t)) return fail(FP_E_ARG, "fp_nl_split_c64: pcoef_s must be 16-byte aligned");
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
  if (p->Nr <= 0) return fail(FP_E_SHAPE, "fp_nl_step_c64: Nr must be positive");
  if (pcoef_misaligned(t)) return fail(FP_E_ARG, "fp_nl_step_c64: pcoef_s must be 16-byte aligned");
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


int fp_batch_plan_create(const fp_batch_instance* inst, int n, fp_batch_plan** plan) {
  if (!inst || !plan || n <= 0) return fail(FP_E_ARG, "fp_batch_plan_create: bad argument");
  *plan = nullptr;
  for (int i = 0; i < n; ++i) {
    const fp_batch_instance& I = inst[i];
    char msg[160];
    if (!valid_nt(I.Nt) || I.Nr <= 0 || I.Nr % 64) {
      snprintf(msg, sizeof msg, "fp_batch_plan_create: instance %d: Nr=%d Nt=%d (need Nr %% 64 == 0, "
               "Nt in {512,1024,2048,4096})", i, I.Nr, I.Nt);
      return fail(FP_E_SHAPE, msg);
    }
    if (I.pp.Nr != I.Nr || I.pp.Nt != I.Nt || I.np.Nr != I.Nr || I.np.Nt != I.Nt ||
        I.pp.enable_ste != I.enable_ste || I.pp.table_size < 2 || I.zc_lo < 0 ||
        I.zc_hi < I.zc_lo || I.zc_hi > 2 * I.Nt) {
      snprintf(msg, sizeof msg, "fp_batch_plan_create: instance %d: parameter blocks disagree with "
               "its shape / STE flag", i);
      return fail(FP_E_ARG, msg);
    }
    if (!I.Yh || !I.Yl || !I.spec || !I.w0 || !I.w1 || !I.rho || !I.phase || !I.mask_r ||
        !I.slots || !I.W_lut || !I.tw || !I.nt.tw || !I.nt.mA || !I.nt.mB || !I.nt.mC ||
        !I.nt.Rf || !I.nt.knots || !I.nt.inv_h || !I.nt.pcoef_s || !I.nt.rw ||
        (I.enable_ste && (!I.rho_s || !I.Ws_lut))) {
      snprintf(msg, sizeof msg, "fp_batch_plan_create: instance %d: null pointer", i);
      return fail(FP_E_ARG, msg);
    }
    if (pcoef_misaligned(&I.nt)) {
      snprintf(msg, sizeof msg, "fp_batch_plan_create: instance %d: pcoef_s must be 16-byte aligned", i);
      return fail(FP_E_ARG, msg);
    }
  }
  fp_batch_plan* P = new fp_batch_plan();
  P->n = n;
  P->host.assign(inst, inst + n);
  if (hipGetDevice(&P->device) != hipSuccess) {
    delete P;
    return fail(FP_E_LAUNCH, "fp_batch_plan_create: no current device");
  }
  if (hipMalloc(&P->dev, sizeof(fp_batch_instance) * n) != hipSuccess ||
      hipMemcpy(P->dev, inst, sizeof(fp_batch_instance) * n, hipMemcpyHostToDevice) != hipSuccess) {
    if (P->dev) (void)hipFree(P->dev);
    delete P;
    return fail(FP_E_LAUNCH, "fp_batch_plan_create: device copy of the descriptors failed");
  }
  P->bm.resize(n);
  P->bn.assign(n, 128);
  P->nl_tpb.resize(n);
  for (int i = 0; i < n; ++i) {
    P->bm[i] = inst[i].Nr % 128 == 0 ? 128 : 64;
    P->nl_tpb[i] = inst[i].Nt >= 4096 ? inst[i].Nt / 16 : inst[i].Nt / 8;
  }
  *plan = P;
  return FP_OK;
}

int fp_batch_plan_destroy(fp_batch_plan* plan) {
  if (!plan) return FP_OK;
  if (plan->dev) (void)hipFree(plan->dev);
  delete plan;
  return FP_OK;
}

int fp_batch_set_config(fp_batch_plan* plan, int i, int bm, int bn, int nl_tpb) {
  if (!plan || i < 0 || i >= plan->n) return fail(FP_E_ARG, "fp_batch_set_config: bad instance");
  const fp_batch_instance& I = plan->host[i];
  const int m = bm ? bm : plan->bm[i], c = bn ? bn : plan->bn[i];
  const int t = nl_tpb ? nl_tpb : plan->nl_tpb[i];
  const bool tile_ok = ((m == 64 || m == 128) && (c == 64 || c == 128)) ||
                       (m == 128 && c == 256) || (m == 256 && c == 128);
  if (!tile_ok || I.Nr % m || (2 * I.Nt) % c)
    return fail(FP_E_SHAPE, "fp_batch_set_config: tile must be 64/128 x 64/128, 128 x 256 or "
                            "256 x 128 and divide Nr x 2 Nt");
  if (t != I.Nt / 8 && t != I.Nt / 4 && !(I.Nt == 4096 && t == I.Nt / 16))
    return fail(FP_E_ARG, "fp_batch_set_config: nonlinear threads must be Nt/8 or Nt/4 "
                          "(or Nt/16 at Nt = 4096)");
  plan->bm[i] = m;
  plan->bn[i] = c;
  plan->nl_tpb[i] = t;
  plan->dirty = true;
  return FP_OK;
}

int fp_batch_set_grid(fp_batch_plan* plan, int kind, int blocks) {
  if (!plan || kind < 0 || kind > 2 || blocks < 0) return fail(FP_E_ARG, "fp_batch_set_grid: bad argument");
  plan->grid[kind] = blocks;
  return FP_OK;
}

int fp_batch_set_merged(fp_batch_plan* plan, int merged) {
  return fp_batch_set_merged_passes(plan, merged ? 0x3F : 0);
}

int fp_batch_set_merged_passes(fp_batch_plan* plan, int mask) {
  if (!plan) return fail(FP_E_ARG, "fp_batch_set_merged_passes: null plan");
  if (mask & ~0x3F) return fail(FP_E_ARG, "fp_batch_set_merged_passes: bits beyond the six passes");
  const int merged = mask;
  if (merged) {
    for (int i = 0; i < plan->n; ++i)
      if (plan->host[i].Nr % 256)
        return fail(FP_E_SHAPE, "fp_batch_set_merged: one launch per pass needs Nr % 256 == 0");
  }
  plan->merged = (unsigned)merged;
  plan->dirty = true;
  return FP_OK;
}

int fp_batch_pass(fp_batch_plan* plan, int pass, void* stream) {
  if (!plan) return fail(FP_E_ARG, "fp_batch_pass: null plan");
  return run_pass(*plan, pass, S(stream));
}

int fp_batch_step_mixed(fp_batch_plan* plan, void* stream) {
  if (!plan) return fail(FP_E_ARG, "fp_batch_step_mixed: null plan");
  static const int order[6] = {FP_PASS_PLASMA, FP_PASS_G1, FP_PASS_G2, FP_PASS_NL, FP_PASS_G3,
                               FP_PASS_G4};
  for (int k = 0; k < 6; ++k) {
    const int rc = run_pass(*plan, order[k], S(stream));
    if (rc) return rc;
  }
  return FP_OK;
}

int fp_batch_launches(const fp_batch_plan* plan) {
  if (!plan) return -1;
  if (plan->dirty) rebuild_groups(const_cast<fp_batch_plan&>(*plan));
  const fp_batch_plan& P = *plan;
  auto on = [&](int pass) { return ((P.merged >> pass) & 1u) != 0; };
  int n = (int)(on(FP_PASS_PLASMA) ? P.plasma_m.size() : P.plasma.size()) +
          (int)(on(FP_PASS_NL) ? P.nl_m.size() : P.nl.size());
  for (int pass : {FP_PASS_G1, FP_PASS_G2, FP_PASS_G3, FP_PASS_G4})
    n += (int)(on(pass) ? P.gemm_m.size() : P.gemm.size());
  return n;
}


int fp_adj_nl_step_c64(const float* v1s, const double* rho, const double* rho_s,
                       const float* g4s, float* g1s, float* grho, float* grho_s, float* scratch,
                       const fp_nl_params* p, const fp_nl_tables_f32* t, unsigned* cmax_out,
                       void* stream) {
  if (!v1s || !rho || !g4s || !g1s || !grho || !scratch || !p || !t)
    return fail(FP_E_ARG, "fp_adj_nl_step_c64: null pointer");
  if (rho_s && !grho_s) return fail(FP_E_ARG, "fp_adj_nl_step_c64: rho_s given without grho_s");
  if (p->Nr <= 0) return fail(FP_E_ARG, "fp_adj_nl_step_c64: bad Nr");
  if (pcoef_misaligned(t)) return fail(FP_E_ARG, "fp_adj_nl_step_c64: pcoef_s must be 16-byte aligned");
  hipStream_t st = S(stream);
  // Nt / 8 threads of 8 samples (222-234 VGPRs, two waves per SIMD, no spills): at
  // 512 x 2048 the backward sweep takes 21 % less time than with Nt / 4 threads of 4
  // samples (six radix-4 passes per transform instead of four radix-8 ones)
#define FP_ADJNL(NT)                                                                           \
  if (rho_s)                                                                                   \
    hipLaunchKernelGGL((fpadj::adj_nl_kernel<NT, NT / 8, true>), dim3(p->Nr), dim3(NT / 8), 0, \
                       st, v1s, rho, rho_s, g4s, g1s, grho, grho_s, scratch, *p, *t, cmax_out); \
  else                                                                                         \
    hipLaunchKernelGGL((fpadj::adj_nl_kernel<NT, NT / 8, false>), dim3(p->Nr), dim3(NT / 8),   \
                       0, st, v1s, rho, nullptr, g4s, g1s, grho, nullptr, scratch, *p, *t,     \
                       cmax_out);                                                              \
  return check_launch("adj_nl_step_c64");
  switch (p->Nt) {
    case 512: { FP_ADJNL(512) }
    case 1024: { FP_ADJNL(1024) }
    case 2048: { FP_ADJNL(2048) }
    case 4096: { FP_ADJNL(4096) }
    default: return fail(FP_E_SHAPE, "fp_adj_nl_step_c64: Nt must be 512, 1024, 2048 or 4096");
  }
#undef FP_ADJNL
}

int fp_adj_plasma_c64(const float* us, const double* rho, const double* rho_s, const float* grho,
                      const float* grho_s, float* gs, const fp_plasma_params* p,
                      const double* W_lut, const double* Ws_lut, const float* tw,
                      unsigned* cmax_out, void* stream) {
  if (!us || !rho || !grho || !gs || !p || !W_lut || !tw)
    return fail(FP_E_ARG, "fp_adj_plasma_c64: null pointer");
  if (p->enable_ste && (!rho_s || !grho_s || !Ws_lut))
    return fail(FP_E_ARG, "fp_adj_plasma_c64: STE on needs rho_s, grho_s and Ws_lut");
  if (p->Nr <= 0 || p->table_size < 2) return fail(FP_E_ARG, "fp_adj_plasma_c64: bad Nr/table_size");
  hipStream_t st = S(stream);
#define FP_ADJPL(NT)                                                                            \
  if (p->enable_ste)                                                                            \
    hipLaunchKernelGGL((fpadj::adj_plasma_ste_kernel<NT, 256>), dim3(p->Nr), dim3(256), 0, st,  \
                       us, rho, rho_s, grho, grho_s, gs, *p, W_lut, Ws_lut, tw, cmax_out);      \
  else                                                                                          \
    hipLaunchKernelGGL((fpadj::adj_plasma_kernel<NT, 256>), dim3(p->Nr), dim3(256), 0, st, us,  \
                       rho, grho, gs, *p, W_lut, tw, cmax_out);                                 \
  return check_launch("adj_plasma_c64");
  switch (p->Nt) {
    case 512: { FP_ADJPL(512) }
    case 1024: { FP_ADJPL(1024) }
    case 2048: { FP_ADJPL(2048) }
    case 4096: { FP_ADJPL(4096) }
    default: return fail(FP_E_SHAPE, "fp_adj_plasma_c64: Nt must be 512, 1024, 2048 or 4096");
  }
#undef FP_ADJPL
}

// ------------------------------------------------------------------ 0-D re-integration
int fp_rate0d_batch(const double* t, const double* I, double* ne, double* ns,
                    const fp_rate0d_params* p, const double* knots, const double* coef_g,
                    const double* coef_s, void* stream) {
  if (!p) return fail(FP_E_ARG, "fp_rate0d_batch: null params");
  if (p->n_rows < 0 || p->Nt < 1 || p->n_knots < 2)
    return fail(FP_E_SHAPE, "fp_rate0d_batch: needs n_rows >= 0, Nt >= 1, n_knots >= 2");
  if (!(p->na > 0.0) || !(p->Imin > 0.0) || !(p->Imax >= p->Imin))
    return fail(FP_E_ARG, "fp_rate0d_batch: needs na > 0 and 0 < Imin <= Imax");
  if (p->n_rows == 0) return FP_OK;   // empty batch: nothing is read or written
  if (!t || !I || !ne || !ns || !knots || !coef_g)
    return fail(FP_E_ARG, "fp_rate0d_batch: null pointer");
  if (p->ste && !coef_s) return fail(FP_E_ARG, "fp_rate0d_batch: STE needs coef_s");
  const unsigned blocks = (unsigned)((p->n_rows + fprate0d::ROWS - 1) / fprate0d::ROWS);
  hipLaunchKernelGGL(fprate0d::rate0d_kernel, dim3(blocks), dim3(fprate0d::ROWS), 0, S(stream), t,
                     I, ne, ns, *p, knots, coef_g, p->ste ? coef_s : coef_g);
  return check_launch("rate0d_batch");
}

// ------------------------------------------------------------------ 1-D SPM + shock
int fp_shock_propagate(const double* E0, double* out, double* work, const fp_shock_row* rows,
                       int n_rows, int N, int nz, int stride, const double* tw, void* stream) {
  if (n_rows < 0 || nz < 0 || stride < 1)
    return fail(FP_E_ARG, "fp_shock_propagate: needs n_rows >= 0, nz >= 0, stride >= 1");
  if (N != 1024 && N != 2048 && N != 4096 && N != 8192)
    return fail(FP_E_SHAPE, "fp_shock_propagate: N must be 1024, 2048, 4096 or 8192");
  if (n_rows == 0) return FP_OK;   // empty batch: nothing is read or written
  if (!E0 || !out || !rows || !tw) return fail(FP_E_ARG, "fp_shock_propagate: null pointer");
  if (N == 8192 && !work) return fail(FP_E_ARG, "fp_shock_propagate: N = 8192 needs work");
  const int n_saves = 1 + nz / stride;
  hipStream_t st = S(stream);
  switch (N) {
    case 1024:
      hipLaunchKernelGGL((fpshock::shock_kernel<1024, 128, false>), dim3(n_rows), dim3(128), 0, st,
                         E0, out, work, rows, nz, stride, n_saves, tw);
      break;
    case 2048:
      hipLaunchKernelGGL((fpshock::shock_kernel<2048, 256, false>), dim3(n_rows), dim3(256), 0, st,
                         E0, out, work, rows, nz, stride, n_saves, tw);
      break;
    case 4096:
      hipLaunchKernelGGL((fpshock::shock_kernel<4096, 512, false>), dim3(n_rows), dim3(512), 0, st,
                         E0, out, work, rows, nz, stride, n_saves, tw);
      break;
    case 8192:
      hipLaunchKernelGGL((fpshock::shock_kernel<8192, 1024, true>), dim3(n_rows), dim3(1024), 0, st,
                         E0, out, work, rows, nz, stride, n_saves, tw);
      break;
    default:
      return fail(FP_E_SHAPE, "fp_shock_propagate: N must be 1024, 2048, 4096 or 8192");
  }
  return check_launch("shock_propagate");
}

}  // extern "C"
