"""Mixed-precision z-step engine (``Config.precision = "mixed"``).

Same split-step algorithm and kernel sequence as the FP64 pipeline of
``integrator.py``, with the precision of each part chosen for MI355X:

* field: complex64, kept in the spectral domain between steps;
* Hankel transform: split-FP16 GEMM on MFMA (``Y = (Yh + Yl)/sY``,
  ``B = (Bh + Bl)/sB``, products ``Yh Bh + Yh Bl + Yl Bh``) with FP32 accumulation;
  the power-of-two scale of B comes from max|B|, left in a device word by the
  kernel that produced B (no host synchronisation);
* propagator phase: evaluated in FP64, reduced mod 2 pi, stored as complex64;
* FFTs, Kerr / Raman / photo-ionisation / plasma terms and RK4: FP32;
* carrier densities: FP64 multiple-shooting scan (the rates span many decades).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import call, ptr, stream_ptr


def _scale_exp(mx):
    """Power-of-two exponent s with mx * 2^s < 2^13 (as the GEMM kernel does for B)."""
    if not (mx > 0) or not np.isfinite(mx):
        return 0
    e = np.frexp(mx)[1]
    return int(13 - e)


def padded_rows(Nr):
    """Row count of the mixed path's GEMM operands: Nr rounded up to a multiple of 128."""
    return -(-int(Nr) // 128) * 128


def split_hankel_fp16(Y):
    """Split Y (float64, Nr x Nr) into FP16 hi/lo at scale 2^s, in the GEMM's fragment
    order [K/32][M/16][4][16][8]: element j of lane (kg, r) is Y[16 mb + r][32 ks + 8 kg + j].
    Y is first zero-padded to Np x Np, Np = Nr rounded up to a multiple of 128: the padded
    rows give zero output rows and the padded columns multiply the zero rows past Nr of the
    field, so the product's first Nr rows are those of the unpadded Y."""
    Nr = Y.shape[0]
    if Y.shape != (Nr, Nr):
        raise ValueError(f"Hankel matrix must be square, got {Y.shape}")
    Np = padded_rows(Nr)
    if Np != Nr:
        Yp = np.zeros((Np, Np))
        Yp[:Nr, :Nr] = Y
        Y = Yp
    M, K = Y.shape
    se = _scale_exp(float(np.abs(Y).max()))
    Ys = Y * 2.0 ** se
    Yh = Ys.astype(np.float16)
    Yl = (Ys - Yh.astype(np.float64)).astype(np.float16)

    def frag(A):
        return np.ascontiguousarray(A.reshape(M // 16, 16, K // 32, 4, 8).transpose(2, 0, 3, 1, 4))
    return frag(Yh), frag(Yl), se


def _exact_pass_table(N, E):
    """Compact twiddles of the exact float transform for rows of E samples per thread
    (fft.hpp, exact_table_offset): per twiddled pass (radix R = min(E, N / Ns) at
    Ns = 1, R, ...; Ns > 1), power-major (r = 1 .. R-1, then k < Ns), the pair (w^r
    rounded to complex64, residual), w = exp(-2 pi i k / (Ns R)), as 4 floats."""
    out, ns = [], 1
    while ns < N:
        R = min(E, N // ns)
        if ns > 1:
            k = np.arange(ns)[None, :]
            r = np.arange(1, R)[:, None]
            z = np.exp(-2j * np.pi * k * r / (ns * R))
            hi = z.astype(np.complex64)
            lo = (z - hi.astype(np.complex128)).astype(np.complex64)
            out.append(np.stack([hi.real, hi.imag, lo.real, lo.imag], axis=-1).reshape(-1, 4))
        ns *= R
    return np.concatenate(out).astype(np.float32) if out else np.zeros((0, 4), np.float32)


def exact_twiddles(N):
    """Twiddle table of the complex64 kernels: tw[m] = exp(-2 pi i m / N) rounded (m < N),
    their residuals tw[N + m] = exp(-2 pi i m / N) - tw[m], then the compact per-pass tables
    of the exact transform for 8, for 4 and for 16 samples per thread (complex64 array)."""
    z = np.exp(-2j * np.pi * np.arange(N) / N)
    hi = z.astype(np.complex64)
    lo = (z - hi.astype(np.complex128)).astype(np.complex64)
    t8, t4, t16 = (_exact_pass_table(N, e) for e in (8, 4, 16))
    flat = np.concatenate([hi.view(np.float32), lo.view(np.float32), t8.ravel(), t4.ravel(),
                           t16.ravel()])
    return np.ascontiguousarray(flat.view(np.complex64))


def matched_scale_pair(x, search=4096):
    """Float32 pair (a, b) with a within `search` ulps of x and a * b as close to 1 as the
    search finds (typically within 1e-10; plain float32(x) * float32(1 / x) is off by up to
    6e-8). The Hankel transform pair scales its two GEMMs by R^2 / j_N and j_N / R^2: in
    float32 their product would be a coherent gain of up to 1.3e-7 per z-step on the whole
    field (5e-4 in amplitude after 4000 steps); a matched pair leaves only the incoherent
    rounding of the products."""
    x32 = np.float32(x)
    ia = np.arange(-search, search + 1, dtype=np.int64) + int(x32.view(np.int32))
    a = ia.astype(np.int32).view(np.float32).astype(np.float64)
    b0 = (1.0 / a).astype(np.float32)
    best = None
    for k in (-1, 0, 1):
        b = np.nextafter(b0, np.float32(np.inf if k > 0 else -np.inf)) if k else b0
        err = np.abs(a * b.astype(np.float64) - 1.0)
        j = int(np.argmin(err))
        if best is None or err[j] < best[0]:
            best = (float(err[j]), float(a[j]), float(b[j]))
    return best[1], best[2]


def zero_columns(spec_mask):
    """Real-column range [lo, hi) of the frequencies where the spectral mask is exactly 0
    (the longest contiguous run; (0, 0) if none): every Hankel GEMM output of the step is
    exactly zero there (the half-step phase carries the mask)."""
    z = np.flatnonzero(np.asarray(spec_mask) == 0.0)
    if z.size == 0:
        return 0, 0
    runs = np.split(z, np.flatnonzero(np.diff(z) != 1) + 1)
    r = max(runs, key=len)
    return 2 * int(r[0]), 2 * (int(r[-1]) + 1)


class MixedTables:
    """Float32 / FP16 device tables of the mixed path, derived from the FP64 ones."""

    def __init__(self, t64, cfg, g):
        dev = t64.device
        self.t64 = t64
        self.Nr, self.Nt = t64.Nr, t64.Nt
        Nt = self.Nt
        c64 = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a, np.complex64))).to(dev)  # noqa: E731
        f32 = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a, np.float32))).to(dev)     # noqa: E731
        self.tw = torch.from_numpy(exact_twiddles(Nt)).to(dev)
        T, iU = g["T_op"], g["inv_U_nl"]
        self.mA, self.mB, self.mC = f32(T * T * iU), f32(-T * iU), f32(iU)
        self.Rf = c64(g["R_f"])
        self.rw = c64(t64.rw_host)
        sp = g["keldysh"]["f_spline"]
        x = sp.x
        h = np.diff(x)
        pp = t64.nl_params(0.0).photo_pref       # same prefactor as the FP64 path
        cs = np.stack([sp.c[0] * h ** 3, sp.c[1] * h ** 2, sp.c[2] * h, sp.c[3]]) * pp
        self.knots = f32(x)
        self.inv_h = f32(1.0 / h)
        self.pcoef_s = f32(cs.T)                  # (n - 1) x 4, one float4 per interval
        yh, yl, se = split_hankel_fp16(g["Y"])
        self.Np = padded_rows(self.Nr)
        self.Yh = torch.from_numpy(yh).to(dev)
        self.Yl = torch.from_numpy(yl).to(dev)
        self.y_scale_exp = se
        mr = np.zeros(self.Np)
        mr[:self.Nr] = g["mask_r_1d"]
        self.mask_r = f32(mr)
        self.Rk, self.iRk = matched_scale_pair(t64.Rk)
        self.zc = zero_columns(g["spec_mask"])
        self._phase = {}
        self.tables = _lib.NLTablesF32(ptr(self.tw), ptr(self.mA), ptr(self.mB), ptr(self.mC),
                                       ptr(self.Rf), ptr(self.knots), ptr(self.inv_h),
                                       ptr(self.pcoef_s), ptr(self.rw))

    def phase_half(self, dz):
        """exp(i phi dz/2) * spec_mask with phi evaluated in FP64 and reduced mod 2 pi
        before rounding to complex64."""
        key = float(dz)
        if key not in self._phase:
            g = self.t64.g
            rho2 = (g["rholist"] ** 2)[:, None]
            ph = (g["delta_k"][None, :] - rho2 / (2 * g["komega"]) * g["inv_U_op"][None, :]) * dz / 2
            ph = np.remainder(ph, 2 * np.pi)
            v = np.zeros((self.Np, self.Nt), np.complex64)      # padded rows: 0
            v[:self.Nr] = np.exp(1j * ph) * g["spec_mask"][None, :]
            self._phase[key] = torch.from_numpy(v).to(self.t64.device)
        return self._phase[key]


class MixedEngine:
    """Spectral-domain complex64 state and the mixed z-step kernel sequence.

    The state and work buffers hold Np = Nr rounded up to a multiple of 128 rows (the
    split-FP16 GEMM's tile height); rows Nr..Np-1 stay zero (the padded Hankel matrix
    writes zeros there), and the row kernels (carrier scan, nonlinear step, record) run
    on the first Nr rows only. ``spec_rows`` is the Nr x Nt view of the state."""

    def __init__(self, t64, cfg, g, u0_c128):
        self.cfg = cfg
        self.t64 = t64
        self.m = MixedTables(t64, cfg, g)
        dev = t64.device
        self.Nr, self.Nt = t64.Nr, t64.Nt
        self.Np = self.m.Np
        self.spec = torch.zeros(self.Np, self.Nt, dtype=torch.complex64, device=dev)
        self.spec_rows = self.spec[:self.Nr]
        self.w0 = torch.zeros_like(self.spec)
        self.w1 = torch.zeros_like(self.spec)
        self.slots = torch.zeros(4, dtype=torch.int32, device=dev)
        self.set_time(u0_c128)

    def _slot(self, i):
        return ctypes.c_void_p(self.slots.data_ptr() + 4 * i)

    def set_time(self, u):
        """Load a time-domain field (any complex dtype) into the spectral state."""
        uc = u.to(torch.complex64).contiguous()
        call("fp_fft_rows_c64", ptr(uc), ptr(self.spec), self.Nr, self.Nt, 0, ptr(self.m.tw),
             stream_ptr())
        self.slots.zero_()
        call("fp_absmax_f32", ptr(self.spec), 2 * self.Np * self.Nt, self._slot(0), stream_ptr())

    def time_c64(self, out=None):
        out = torch.empty_like(self.spec_rows) if out is None else out
        call("fp_fft_rows_c64", ptr(self.spec), ptr(out), self.Nr, self.Nt, 1, ptr(self.m.tw),
             stream_ptr())
        return out

    def plasma(self, rho, rho_s):
        t = self.t64
        call("fp_plasma_scan_spec_c64", ptr(self.spec), ptr(rho),
             ptr(rho_s) if self.cfg.enable_ste else None, None, ctypes.byref(t.plasma_params),
             ptr(t.W_lut), ptr(t.Ws_lut), ptr(self.m.tw), self._slot(1), ptr(t.status),
             stream_ptr())

    def strang(self, dz, rho, rho_s_arg, nl_params, clear_slots=True):
        """Strang step on the spectral state. Slots 1-3 must start at zero: the plasma
        kernel clears them in the pipeline, otherwise they are cleared here."""
        m, st = self.m, stream_ptr()
        if clear_slots:
            self.slots[1:].zero_()
        Nr, Nt = self.Np, self.Nt          # the GEMMs run over the padded rows
        ph = m.phase_half(dz)
        zlo, zhi = m.zc
        # G1 / G3 leave the masked columns of w1 unwritten (G2 / G4 do not read them); G2 /
        # G4 write zeros there, since the nonlinear step and the carrier scan read whole rows
        g = lambda B, C, a, rs, phase, bin_, cout, reset, zf: call(  # noqa: E731
            "fp_hankel_gemm_mixed_band", ptr(m.Yh), ptr(m.Yl), m.y_scale_exp, ptr(B), ptr(C), Nr,
            Nt, ctypes.c_float(a), ptr(rs), ptr(phase), bin_, cout, reset, zlo, zhi, zf, st)
        g(self.spec, self.w1, m.Rk, None, ph, self._slot(0), self._slot(1), None, 0)
        g(self.w1, self.w0, m.iRk, None, None, self._slot(1), None, self._slot(0), 1)
        call("fp_nl_step_c64", ptr(self.w0), ptr(self.w0), ptr(rho), ptr(rho_s_arg), 1, 1,
             ctypes.byref(nl_params), ctypes.byref(m.tables), self._slot(2), st)
        g(self.w0, self.w1, m.Rk, None, ph, self._slot(2), self._slot(3), None, 0)
        g(self.w1, self.spec, m.iRk, m.mask_r, None, self._slot(3), self._slot(0), None, 1)

    def step(self, dz, rho, rho_s, rho_s_arg, nl_params):
        self.plasma(rho, rho_s)
        self.strang(dz, rho, rho_s_arg, nl_params, clear_slots=False)
