/*
 * rapport_amd -- C ABI of the MI355X (gfx950) split-step filamentation propagator.
 *
 * Plain C: device pointers, sizes and POD parameter blocks; no torch types.
 * Every device pointer is caller-owned memory on the current HIP device;
 * `stream` is a hipStream_t (NULL = default stream). Calls are asynchronous on
 * that stream. Return value: 0 on success, otherwise an FP_E_* code or a
 * hipError_t value; fp_last_error() gives a message.
 *
 * Layouts: fields are row-major (r, t) with interleaved complex (re, im);
 * densities are row-major (r, t) float64; "spectral" rows hold the numpy-
 * convention FFT along t of a field row (frequency order of numpy.fft.fftfreq).
 *
 * Reference citations are relative to the reference repository;
 * D3 = ".github/deliverable (3)".
 */
#ifndef RAPPORT_AMD_H
#define RAPPORT_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define FP_OK 0
#define FP_E_ARG 9001      /* invalid argument (shape, null pointer, flag) */
#define FP_E_SHAPE 9002    /* Nt / Nr not supported by the compiled kernels */
#define FP_E_LAUNCH 9003   /* kernel launch failed */

/* Scalars of the carrier rate kernel, one-to-one with the arguments of
 * solve_rate_equation_kernel (D3/sim/kernels.py:36-46). */
typedef struct fp_plasma_params {
  int Nr, Nt;
  double dt;
  double field_to_I;   /* invE2 */
  double beta_g, beta_s;
  double na;           /* rho_max */
  double inv_tau_r, inv_tau_ste;
  int enable_ste;
  int table_size;
  double log_min_I, inv_log_step, Imin, Imax;
} fp_plasma_params;

/* Scalars of NonlinearOperator.split (D3/sim/operators.py:59-146) and of the
 * nonlinear part of Integrator.step (D3/sim/integrator.py:141-147). */
typedef struct fp_nl_params {
  int Nr, Nt;
  double dz;
  double invE2, Imin, Imax;
  double photo_pref;      /* 1e6 * Ui / (n0 c eps0): photo = W * photo_pref / (|u|^2 + 1e-30) * depl */
  double rho_max;
  double plasma_pref;     /* sigma/2 * 100 */
  double pc_re, pc_im;    /* plasma_coeff (absorption + i omega0 tau_c defocusing, per toggles) */
  double ste_pref;        /* 0 disables the STE index term */
  double kerr_pref, f_R;
  int en_kerr_inst, en_kerr_raman, en_pi_loss, en_plasma_absorb;
  int n_knots;            /* PCHIP knots: 17 decades x 900, 10^k (1 + 0.01 j) */
} fp_nl_params;

/* Device tables shared by the nonlinear kernels (float64). */
typedef struct fp_nl_tables {
  const double* tw;       /* Nt complex: exp(-2 pi i m / Nt) */
  const double* mA;       /* Nt: T^2 * invU_nl  (Kerr bracket multiplier)     */
  const double* mB;       /* Nt: -T * invU_nl   (photo-ionisation multiplier) */
  const double* mC;       /* Nt: invU_nl        (plasma / STE multiplier)     */
  const double* Rf;       /* Nt complex: Raman transfer function R_f          */
  const double* knots;    /* n_knots PCHIP abscissae (W/cm^2)                 */
  const double* pcoef;    /* 4 x (n_knots-1) PCHIP power-basis coefficients   */
  const double* rw;       /* Nt+2 complex: w^m (m = 0..Nt), then (C, 0); the Raman
                             response as a recursion, w = exp(dt (-1/tau_d + i/tau_s)),
                             C = dt Omega_r2 tau_s (see raman.hpp)                 */
} fp_nl_tables;

/* Same tables in float32 for the mixed-precision path. PCHIP coefficients are
 * in the local coordinate s = (I - x_i)/h_i and pre-scaled by photo_pref, so
 * photo = |P(s)| * depl / (|u|^2 + 1e-30) stays inside float range. */
typedef struct fp_nl_tables_f32 {
  const float* tw;        /* complex64: twiddles tw[m] = exp(-2 pi i m / Nt) rounded (m < Nt),
                             their residuals tw[Nt + m] = exp(-2 pi i m / Nt) - tw[m], then
                             from tw[2 Nt] the compact per-pass tables of the exact
                             transforms (rows of 8, 4 and 16 samples per thread; per pass,
                             power-major, the pair (w^r rounded, residual) as 4 floats), read
                             by the nonlinear step's transforms into and out of the row, which
                             have no coherent rounding gain; rapport_amd.mixed.exact_twiddles
                             builds it */
  const float* mA;
  const float* mB;
  const float* mC;
  const float* Rf;        /* Nt complex64 */
  const float* knots;     /* n_knots */
  const float* inv_h;     /* n_knots-1: 1 / (x_{i+1} - x_i) */
  const float* pcoef_s;   /* (n_knots-1) x 4 interleaved (c0..c3 per interval), 16-byte aligned */
  const float* rw;        /* Nt+2 complex64: as fp_nl_tables.rw */
} fp_nl_tables_f32;

int fp_version(void);
const char* fp_last_error(void);

/* ---------------------------------------------------------------------------
 * Carrier rate equations along t for all rows.
 * Replaces NonlinearOperator.update_plasma + rate_eq_kernel launch
 * (D3/sim/operators.py:165-183; kernel D3/sim/kernels.py:36-106).
 * u: Nr x Nt complex128 (time domain). rho, rho_s: Nr x Nt float64, fully
 * written (column 0 = 0 as update_plasma sets it). rho_s may be NULL when
 * enable_ste == 0. status (optional, device int): incremented per row whose
 * Newton scan hit its iteration cap. Supported Nt: 512, 1024, 2048, 4096.
 */
int fp_plasma_scan(const double* u, double* rho, double* rho_s, const fp_plasma_params* p,
                   const double* W_lut, const double* Ws_lut, int* status, void* stream);

/* Same, fused with an inverse FFT: `spec` holds the spectral rows of the field
 * (Nr x Nt complex128); the time-domain field is formed in LDS and never
 * written. Used by the z-step pipeline (Integrator.propagate,
 * D3/sim/integrator.py:157). If u_out is non-NULL the time-domain field is
 * also stored there. */
int fp_plasma_scan_spec(const double* spec, double* rho, double* rho_s, double* u_out,
                        const fp_plasma_params* p, const double* W_lut, const double* Ws_lut,
                        const double* tw, int* status, void* stream);

/* ---------------------------------------------------------------------------
 * Batched FFT along t of Nr rows of Nt complex128 (numpy.fft.fft / ifft along
 * axis=1, D3/sim/operators.py:49,52). inverse=1 applies the 1/Nt scale.
 * in may equal out. */
int fp_fft_rows(const double* in, double* out, int Nr, int Nt, int inverse, const double* tw,
                void* stream);

/* ---------------------------------------------------------------------------
 * Hankel GEMM on FP64 MFMA: C = diag(rowscale) * alpha * (Y @ B), optionally
 * followed by the complex multiply C[m, f] *= phase[m, f].
 * Replaces cp.dot(self.Y, u) (D3/sim/operators.py:48,53).
 * Yt: Y transposed (Nr x Nr, row-major Yt[k][m] = Y[m][k]).
 * B, C: Nr x Nt complex128 viewed as Nr x 2Nt real. rowscale (Nr) and phase
 * (Nr x Nt complex128) may be NULL. C must not alias B. */
int fp_hankel_gemm(const double* Yt, const double* B, double* C, int Nr, int Nt, double alpha,
                   const double* rowscale, const double* phase, void* stream);

/* ---------------------------------------------------------------------------
 * LinearOperator.half_linear (D3/sim/operators.py:47-53) on time-domain rows:
 * out = ifft_t(iRk * Y @ (phase_half o (Rk * Y @ fft_t(u)))), which equals the
 * reference's iRk Y ifft(phase fft(Rk Y u)) because Y (along r) commutes with
 * the FFT (along t). phase_half[m, f] = exp(i phi[m, f] dz/2) * spec_mask[f]
 * (Nr x Nt complex128). work: 2 x Nr x Nt complex128 scratch. u may equal out. */
int fp_half_linear(const double* u, double* out, int Nr, int Nt, const double* Yt,
                   const double* phase_half, double Rk, double iRk, const double* tw,
                   double* work, void* stream);

/* ---------------------------------------------------------------------------
 * NonlinearOperator.split (D3/sim/operators.py:89-146): rhs (Nr x Nt
 * complex128) and alpha (Nr x Nt float64, may be NULL) for a time-domain field.
 * rho_s may be NULL (no STE index term). */
int fp_nl_split(const double* u, const double* rho, const double* rho_s, double* rhs,
                double* alpha, const fp_nl_params* p, const fp_nl_tables* t, void* stream);

/* ---------------------------------------------------------------------------
 * Nonlinear part of Integrator.step (D3/sim/integrator.py:141-147) fused per
 * row: alpha half-step, classical RK4 over split() with rho frozen, alpha
 * half-step. in_spectral / out_spectral select spectral or time-domain rows
 * for input and output (the z-step pipeline uses spectral in and out).
 * in may equal out. */
int fp_nl_step(const double* in, double* out, const double* rho, const double* rho_s,
               int in_spectral, int out_spectral, const fp_nl_params* p, const fp_nl_tables* t,
               void* stream);

/* ---------------------------------------------------------------------------
 * Record reductions of Integrator._record / NonlinearOperator.loss_rates
 * (D3/sim/integrator.py:182-212; D3/sim/operators.py:148-163) for the
 * time-domain field u and densities rho, rho_s (rho_s may be NULL).
 * rows_out: 7 x Nr float64 = [sum_t |u|^2, max_t rho, max_t rho_s,
 *   max_t I, sum_t dep, sum_t 2 photo I, sum_t 2 plasma I] where
 * I = |u|^2 invE2 and dep = 2 (photo + plasma) I dt 100.
 * bins (optional, nbins x Nr): sum_t dep over t-index ranges
 * [bin_lo[b], bin_hi[b]). */
int fp_record_rows(const double* u, const double* rho, const double* rho_s, double* rows_out,
                   const fp_nl_params* p, const fp_nl_tables* t, double dt, int nbins,
                   const int* bin_lo, const int* bin_hi, double* bins, void* stream);

/* Destinations of one record of Integrator._record (D3/sim/integrator.py:179-224):
 * the entries at the record's index of each history, float32 unless noted.
 * Row vectors hold Nr values, on-axis rows Nt, the cube Nr_sub x Nt_sub with
 * Nr_sub = (Nr-1)/r_stride + 1, Nt_sub = (Nt-1)/t_stride + 1; bins is nbins x Nr. */
typedef struct fp_record_dst {
  float* fluence;       /* sum_t |u|^2 dt invE2 */
  float* rho;           /* max_t rho */
  float* rho_s;         /* max_t rho_s */
  float* ipeak;         /* max_t I */
  float* absorbed;      /* sum_t dep */
  float* bins;          /* sum_t dep over [bin_lo[b], bin_hi[b]); NULL iff nbins == 0 */
  float* rho_at;        /* rho[:, it_snap]; NULL to skip */
  int it_snap;
  float* rho_onaxis;    /* rho[0, :] */
  float* rho_s_onaxis;  /* rho_s[0, :] */
  float* I_onaxis;      /* I[0, :] */
  int r_stride, t_stride, Nt_sub;  /* t_stride = 0 skips the cube */
  float* rho_cube;      /* rho[::r_stride, ::t_stride] */
  float* rho_s_cube;
  float* I_cube;
  const double* r_w;    /* Nr radial quadrature weights 2 pi r dr */
  double* partial;      /* 3 x Nr FP64 scratch */
  double* imax;         /* FP64: max_r max_t I */
  double* dedz_photo;   /* FP64: sum_r (sum_t 2 photo I) dt r_w 1e10 */
  double* dedz_plasma;  /* FP64: sum_r (sum_t 2 plasma I) dt r_w 1e10 */
} fp_record_dst;

/* One whole record from the spectral state `spec` (Nr x Nt, complex128 when
 * spec_c64 == 0 with the FP64 twiddles `tw`, else complex64 with the FP32
 * twiddles): the inverse transform along t and the reductions of
 * fp_record_rows, written straight into the histories of `d`. Replaces the
 * inverse fp_fft_rows + fp_record_rows + per-history copies of
 * Integrator._record (D3/sim/integrator.py:179-224). rho_s may be NULL (read
 * as zero). */
int fp_record_history(const void* spec, int spec_c64, const void* tw, const double* rho,
                      const double* rho_s, const fp_nl_params* p, const fp_nl_tables* t,
                      double dt, int nbins, const int* bin_lo, const int* bin_hi,
                      const fp_record_dst* d, void* stream);

/* Largest Nr that fp_record_history_noalloc accepts. */
#define FP_NOALLOC_MAX_NR 16384

/* fp_record_history without caller scratch: the same kernels and results, with the
 * 3 x Nr per-row partials in a static device array of the library (3 x
 * FP_NOALLOC_MAX_NR doubles per device, kept until the process ends); d->partial is
 * ignored. Has no counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops). It calls
 * no allocator. Returns FP_E_SHAPE when Nr > FP_NOALLOC_MAX_NR. Uses the array of the
 * current device, so `stream` must belong to that device. Every call on one device
 * shares the array: calls that overlap on one device (concurrent streams or host
 * threads) race; calls ordered on one stream do not (every partial is written before
 * it is read). */
int fp_record_history_noalloc(const void* spec, int spec_c64, const void* tw, const double* rho,
                              const double* rho_s, const fp_nl_params* p, const fp_nl_tables* t,
                              double dt, int nbins, const int* bin_lo, const int* bin_hi,
                              const fp_record_dst* d, void* stream);

/* ===========================================================================
 * Mixed-precision path (Config.precision = "mixed"): complex64 fields, FP32
 * transforms and RK4, split-FP16 Hankel GEMM with FP32 accumulation, FP64
 * carrier densities. Same operations and layouts as the FP64 entry points.
 * "slots" are device words holding max|x| of a buffer as float bits (written
 * with atomicMax, cleared by the kernels named below), used to pick the
 * power-of-two FP16 scale of the next GEMM without a host round trip.
 * ======================================================================== */

/* max(|re|, |im|) over n floats into *slot (atomicMax of float bits). */
int fp_absmax_f32(const float* x, long long n, unsigned* slot, void* stream);

/* Batched complex64 FFT along t (cf. fp_fft_rows). */
int fp_fft_rows_c64(const float* in, float* out, int Nr, int Nt, int inverse, const float* tw,
                    void* stream);

/* Carrier rate equations (FP64) from complex64 spectral rows (cf.
 * fp_plasma_scan_spec; replaces update_plasma, D3/sim/operators.py:165-183).
 * clear_slots (optional): 3 consecutive slot words set to 0 by the kernel. */
int fp_plasma_scan_spec_c64(const float* spec, double* rho, double* rho_s, float* u_out,
                            const fp_plasma_params* p, const double* W_lut, const double* Ws_lut,
                            const float* tw, unsigned* clear_slots, int* status, void* stream);

/* Hankel GEMM, split FP16 x 3 on MFMA (replaces cp.dot(self.Y, u),
 * D3/sim/operators.py:48,53). Yh/Yl: Y * 2^y_scale_exp split into FP16 hi/lo in
 * fragment order (see rapport_amd/mixed.py). B, C: Nr x Nt complex64. bmax_in:
 * slot holding max|B|; cmax_out (optional): slot receiving max|C|; reset_slot
 * (optional): cleared by the kernel before any of its blocks can write it.
 * Requires Nr % 128 == 0 and (2 Nt) % 128 == 0. */
int fp_hankel_gemm_mixed(const void* Yh, const void* Yl, int y_scale_exp, const float* B,
                         float* C, int Nr, int Nt, float alpha, const float* rowscale,
                         const float* phase, const unsigned* bmax_in, unsigned* cmax_out,
                         unsigned* reset_slot, void* stream);

/* fp_hankel_gemm restricted to the columns outside [zc_lo, zc_hi) (real columns), where
 * the caller knows the product is exactly zero (the spectral mask is 0 there): the other
 * columns are tiled to fill the GPU once and are bitwise fp_hankel_gemm's values; columns in
 * the range are not written, or, with zero_fill, set to 0. An empty or too narrow range
 * (zc_lo = zc_hi = 0) tiles all columns the same way when 128 x 128 tiles would not fill
 * the GPU. Falls back to fp_hankel_gemm otherwise, or when Nr % 128 != 0. Replaces
 * cp.dot(self.Y, u) (D3/sim/operators.py:48,53) in the z-step pipeline. */
int fp_hankel_gemm_band(const double* Yt, const double* B, double* C, int Nr, int Nt,
                        double alpha, const double* rowscale, const double* phase, int zc_lo,
                        int zc_hi, int zero_fill, void* stream);

/* fp_hankel_gemm_mixed restricted to the columns outside [zc_lo, zc_hi) (real columns,
 * 0 <= zc_lo <= zc_hi <= 2 Nt), where the caller knows the product is exactly zero: the
 * spectral mask is 0 there, so the G1 / G3 outputs (phase x mask) and the G2 / G4 inputs
 * vanish. The remaining columns are tiled so that the tiles fill the GPU once; columns in
 * the range are not written, or, with zero_fill, set to 0 (needed where a later row
 * kernel reads the full row). With a phase, the same results as fp_hankel_gemm_mixed in the
 * columns computed; without one (G2 / G4 of the z-step) the tile splits K between two wave
 * groups, so the results agree with fp_hankel_gemm_mixed to FP32 accumulation rounding (the
 * products are the same, the summation order differs). When the range is too narrow to gain
 * anything (or Nr % 256 != 0) it is fp_hankel_gemm_mixed. Replaces cp.dot(self.Y, u)
 * (D3/sim/operators.py:48,53) in the z-step pipeline. */
int fp_hankel_gemm_mixed_band(const void* Yh, const void* Yl, int y_scale_exp, const float* B,
                              float* C, int Nr, int Nt, float alpha, const float* rowscale,
                              const float* phase, const unsigned* bmax_in, unsigned* cmax_out,
                              unsigned* reset_slot, int zc_lo, int zc_hi, int zero_fill,
                              void* stream);

/* split() on complex64 rows (cf. fp_nl_split). */
int fp_nl_split_c64(const float* u, const double* rho, const double* rho_s, float* rhs,
                    float* alpha, const fp_nl_params* p, const fp_nl_tables_f32* t,
                    void* stream);

/* Fused nonlinear step on complex64 rows (cf. fp_nl_step); cmax_out optional. */
int fp_nl_step_c64(const float* in, float* out, const double* rho, const double* rho_s,
                   int in_spectral, int out_spectral, const fp_nl_params* p,
                   const fp_nl_tables_f32* t, unsigned* cmax_out, void* stream);

/* ---------------------------------------------------------------- adjoint
 * No counterpart in the reference (it has no gradients): the discrete adjoint of
 * one mixed-precision z-step, with or without STE. Gradients use the convention
 * dL = Re sum conj(g) du and travel between calls as FFT_t(g) ("spectral"), like
 * the field state. With the split-FP16 GEMM applied to Y^T (and Y^T diag(mask_r))
 * and the conjugate half-step phase, one adjoint step is
 *   A1 = iRk (Y^T diag(mask_r) gamma) x conj(phase), g4 = Rk Y^T A1,
 *   fp_adj_nl_step_c64 (g4 -> g1, dL/drho, dL/drho_s),
 *   A3 = iRk (Y^T g1) x conj(phase), gamma_n = Rk Y^T A3,
 *   fp_adj_plasma_c64 (gamma_n += FFT_t(2 invE2 dL/dI u_n)). */

/* Adjoint of the nonlinear half of Integrator.step (integrator.py:142-146) at the
 * state v1s (spectral, the field after the first half linear step; rho = the
 * carrier densities of the step). rho_s: the step's STE density when the forward
 * step applied the STE index term (the rho_s argument of fp_nl_step_c64), else
 * NULL; grho_s (required with rho_s): float Nr x Nt dL/drho_s. g4s: spectral
 * gradient w.r.t. the input of the second half linear step; g1s (may alias g4s):
 * spectral gradient w.r.t. v1s; grho: float Nr x Nt dL/drho; scratch: 4 Nr x Nt
 * complex64 planes; cmax_out (optional): max |re|, |im| of g1s (float bits,
 * atomicMax). */
int fp_adj_nl_step_c64(const float* v1s, const double* rho, const double* rho_s,
                       const float* g4s, float* g1s, float* grho, float* grho_s, float* scratch,
                       const fp_nl_params* p, const fp_nl_tables_f32* t, unsigned* cmax_out,
                       void* stream);

/* Adjoint of the carrier scan (fp_plasma_scan_spec_c64, kernels.py:36-106):
 * us = the spectral state the scan read, rho (and, with p->enable_ste, rho_s) = its
 * result, grho (grho_s) = dL/drho (dL/drho_s); W_lut, Ws_lut: the scan's rate
 * tables (rho_s, grho_s, Ws_lut may be NULL with STE off); adds
 * FFT_t(2 invE2 (dL/dI) u) to the spectral gradient gs; cmax_out (optional): max of
 * the updated gs. */
int fp_adj_plasma_c64(const float* us, const double* rho, const double* rho_s, const float* grho,
                      const float* grho_s, float* gs, const fp_plasma_params* p,
                      const double* W_lut, const double* Ws_lut, const float* tw,
                      unsigned* cmax_out, void* stream);

/* ---------------------------------------------------------------- batched form
 * No counterpart in the reference, which runs one propagation per process: many
 * independent mixed-precision propagations ("instances", each with its own Nr x Nt,
 * pulse and material parameters) advance together, one grouped launch per pass.
 * The per-instance semantics are exactly those of the single-instance mixed
 * sequence: fp_plasma_scan_spec_c64 (clears slots 1-3) -> G1 fp_hankel_gemm_mixed
 * (spec -> w1, Rk, phase, slot 0 -> slot 1) -> G2 (w1 -> w0, iRk, slot 1, resets
 * slot 0) -> fp_nl_step_c64 (w0 in place, spectral in/out, max -> slot 2) -> G3
 * (w0 -> w1, Rk, phase, slot 2 -> slot 3) -> G4 (w1 -> spec, iRk, mask_r, slot 3 ->
 * slot 0). All pointers are device pointers owned by the caller. */
typedef struct fp_batch_instance {
  int Nr, Nt;                 /* Nr % 64 == 0; Nt in {512, 1024, 2048, 4096} */
  int y_scale_exp;            /* Y = (Yh + Yl) 2^-y_scale_exp */
  int enable_ste;             /* 1: rho_s is evolved and enters the nonlinear step */
  float Rk, iRk;              /* R^2/j_N and j_N/R^2 */
  const void* Yh;             /* split Hankel matrix, fragment order (fp_hankel_gemm_mixed) */
  const void* Yl;
  float* spec;                /* Nr x Nt complex64 state: FFT_t of the field */
  float* w0;                  /* Nr x Nt complex64 work buffers */
  float* w1;
  double* rho;                /* Nr x Nt carrier densities */
  double* rho_s;              /* (required when enable_ste) */
  const float* phase;         /* Nr x Nt complex64 half-step phase x spectral mask */
  const float* mask_r;        /* Nr radial absorbing mask */
  unsigned* slots;            /* 4 scale words (float bits); slot 0 = max|spec| */
  const double* W_lut;        /* fp_plasma_scan tables */
  const double* Ws_lut;
  const float* tw;            /* Nt complex64 twiddles exp(-2 pi i k / Nt) */
  int* status;                /* plasma-scan cap counter */
  fp_plasma_params pp;
  fp_nl_params np;
  fp_nl_tables_f32 nt;
  int zc_lo, zc_hi;           /* real columns [zc_lo, zc_hi) of every GEMM output are exactly
                                 zero (frequencies where the spectral mask is 0): those
                                 tiles are written as zeros without the product; 0, 0 = none */
} fp_batch_instance;

typedef struct fp_batch_plan fp_batch_plan;   /* opaque */

/* Passes of one batched z-step, in order (fp_batch_pass). */
enum { FP_PASS_PLASMA = 0, FP_PASS_G1 = 1, FP_PASS_G2 = 2, FP_PASS_NL = 3, FP_PASS_G3 = 4,
       FP_PASS_G4 = 5 };

/* Validates the n descriptors (host array), copies them to the device and groups
 * the instances per pass. Default launch configuration: GEMM tile 128 x 128 (64
 * rows when Nr % 128 != 0), nonlinear kernel Nt/8 threads (Nt/16 at Nt = 4096), one block
 * per work item. */
int fp_batch_plan_create(const fp_batch_instance* inst, int n, fp_batch_plan** plan);
int fp_batch_plan_destroy(fp_batch_plan* plan);

/* Launch configuration of instance i: GEMM tile bm x bn (64/128 x 64/128 with 4 waves,
 * 128 x 256 with 4 waves, or 256 x 128 with 8 waves; Nr % bm == 0, 2 Nt % bn == 0) and
 * nonlinear threads per row nl_tpb (Nt/8 or Nt/4, or Nt/16 at Nt = 4096); 0 keeps the
 * current value. */
int fp_batch_set_config(fp_batch_plan* plan, int i, int bm, int bn, int nl_tpb);

/* Grid of the launches of pass kind 0 = plasma, 1 = GEMM, 2 = nonlinear: 0 = one
 * block per row (tile); > 0 = that many persistent blocks, each looping over work items. */
int fp_batch_set_grid(fp_batch_plan* plan, int kind, int blocks);

/* merged = 1: every pass is one launch for the whole batch (all Nt classes, STE flags and
 * tile shapes in one work list; one block per work item: a row of the carrier scan, 4096 / Nt
 * rows of the nonlinear step at Nt / 8 threads per row, a 256 x 128 GEMM tile), so a z-step
 * is six launches; the per-instance tiles, threads and grids of fp_batch_set_config /
 * fp_batch_set_grid then do not apply. Needs Nr % 256 == 0 (FP_E_SHAPE otherwise, and the
 * plan keeps its launch mode). merged = 0: one launch per
 * group of instances sharing a kernel configuration (the default). */
int fp_batch_set_merged(fp_batch_plan* plan, int merged);

/* Launch mode per pass: bit FP_PASS_x of mask set = that pass is one launch over the merged
 * work list of every instance (as fp_batch_set_merged(plan, 1) does for all six), clear = the
 * grouped launches. fp_batch_set_merged(plan, m) is this call with mask 0x3F (m != 0) or 0.
 * Needs Nr % 256 == 0 when mask != 0 (FP_E_SHAPE otherwise, the plan keeps its mode).
 * No counterpart in the reference (the batched form is new). */
int fp_batch_set_merged_passes(fp_batch_plan* plan, int mask);

/* One pass (FP_PASS_*) for every instance, or the whole z-step (fp_batch_step_mixed). */
int fp_batch_pass(fp_batch_plan* plan, int pass, void* stream);
int fp_batch_step_mixed(fp_batch_plan* plan, void* stream);

/* Number of kernel launches one fp_batch_step_mixed makes (after grouping). */
int fp_batch_launches(const fp_batch_plan* plan);

/* ================================================================ 0-D rate re-integration
 * Batched replacement of integre_0d (D3/sim/figures_article.py:168-194, with _phi1 / _step
 * :161-166): the electron and STE densities re-integrated over given intensity histories
 * I(t), one independent row per history (fig13_electron_density_vs_z calls it once per
 * saved z plane and model variant, :297-303). Per step k -> k+1: averages of I, W, W_s
 * over the two samples, the exact exponential step n <- e^{L dt} n + S dt phi1(L dt)
 * (0 when not finite or negative), then n_e, n_s rescaled when n_e + n_s > rho_max. */
typedef struct fp_rate0d_params {
  int n_rows, Nt;                 /* intensity histories x samples per history (Nt >= 1) */
  double na;                      /* rho_max [cm^-3] */
  double inv_tau_r;               /* 1 / tau_r [1/s] */
  double beta_g, beta_s;          /* sigma / Ui, sigma / Us [cm^2/J] */
  int avalanche, trapping, ste;   /* model variants (avalanche, piegeage, ste) */
  int n_knots;                    /* PCHIP knots in log10 I (>= 2) */
  double Imin, Imax;              /* clip range of I for the rates [W/cm^2] */
} fp_rate0d_params;

/* t: Nt sample times [s], shared by every row; I: n_rows x Nt intensities [W/cm^2];
 * knots: n_knots increasing log10 I knots; coef_g, coef_s: PCHIP coefficients of the
 * valence-band and STE rates, 4 x (n_knots - 1), row j multiplying (x - knot_i)^(3 - j)
 * (scipy PPoly order); coef_s is read only when ste. Rate = max(pchip(log10(clip(I))), 0).
 * ne, ns: n_rows x Nt outputs (ns holds zeros when ste == 0). All device pointers. */
int fp_rate0d_batch(const double* t, const double* I, double* ne, double* ns,
                    const fp_rate0d_params* p, const double* knots, const double* coef_g,
                    const double* coef_s, void* stream);

/* ================================================================ 1-D SPM + shock model
 * Batched replacement of propagate_shock (D3/sim/self_steepening.py:80-105, right-hand
 * side shock_rhs :65-77 and _dt_spectral :60-62): classical RK4 in z of
 *   dE/dz = i spm |E|^2 E - shock ifft(i w fft(|E|^2 E)),  w = 2 pi fftfreq(N, dt),
 * one workgroup per pulse for the whole z march (field in registers, transforms in LDS). */
typedef struct fp_shock_row {
  double dt;      /* sample spacing [s] */
  double dz;      /* z step [m] */
  double spm;     /* K0 n2 (E in sqrt(W/m^2)) */
  double shock;   /* n2 / c; 0 = SPM only */
} fp_shock_row;

/* E0: n_rows x N complex128 input pulses; rows: n_rows parameter blocks; out: n_rows x
 * n_saves x N complex128 with plane 0 = E0 and plane j the field after j * stride steps,
 * n_saves = 1 + nz / stride; work: n_rows x 3N complex128 scratch, used for N = 8192 (may be
 * NULL otherwise); tw: N complex128 twiddles exp(-2 pi i m / N). N in {1024, 2048, 4096,
 * 8192}. All device pointers. */
int fp_shock_propagate(const double* E0, double* out, double* work, const fp_shock_row* rows,
                       int n_rows, int N, int nz, int stride, const double* tw, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* RAPPORT_AMD_H */
