"""Differentiation form: hand-written adjoint of the mixed-precision z-step.

``AdjointEngine`` wraps a mixed-precision ``Integrator`` and computes the
output-plane fluence F(r) = sum_t |u(r, t)|^2 dt invE2 after n z-steps (the
propagate() loop body: carrier scan + Strang step) and the gradient of a loss of F
with respect to the input field, by the discrete adjoint of every step:

    gamma <- A1 = iRk (Y^T diag(mask_r) gamma) x conj(phase)   [split-FP16 GEMM]
             g4 = Rk Y^T A1                                     [split-FP16 GEMM]
             fp_adj_nl_step_c64: g4 -> g1 (+ dL/drho, dL/drho_s) [row kernel]
             A3 = iRk (Y^T g1) x conj(phase); gamma = Rk Y^T A3 [GEMMs]
             fp_adj_plasma_c64: gamma += FFT_t(2 invE2 dL/dI u)  [row kernel]

With STE on (``enable_ste``), the carrier scan's adjoint runs over the coupled
(n_e, n_s) state and the nonlinear step's adjoint differentiates the STE index term
i ste_pref rho_s u as well (dL/drho_s feeds the scan's adjoint).

Gradients travel as FFT_t of the real-gradient field g (dL = Re sum conj(g) du),
the same representation as the spectral field state. The forward states are
checkpointed every ``ckpt`` steps (default ceil(sqrt(n))) and each segment is
recomputed once during the backward sweep, keeping per step the state, its carrier
densities and the field after the first half linear step (what the step's adjoint
linearises around), so memory holds about sqrt(n) states + 3 sqrt(n) step records
(4 with STE: rho_s too).

``FluenceFromInput`` exposes the forward / backward pair as a torch.autograd.Function
(input: complex64 time-domain field, output: F as float64). ``lg_modes`` /
``lg_field`` build the input from a radial Laguerre-Gauss amplitude + phase
expansion and ``fit_lg`` fits its coefficients to a target fluence.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from ._lib import call, ptr, stream_ptr
from .mixed import split_hankel_fp16


class AdjointEngine:
    def __init__(self, integ, ckpt=None):
        if integ._mx is None:
            raise ValueError("AdjointEngine needs a mixed-precision Integrator (precision='mixed')")
        if integ._mx.Np != integ.Nr:
            raise ValueError(f"the adjoint needs Nr % 128 == 0 (Nr = {integ.Nr})")
        self.it = integ
        self.mx = integ._mx
        self.ckpt = ckpt
        g, dev = integ.g, integ.device
        self.Nr, self.Nt, self.dz = integ.Nr, integ.Nt, integ.dz
        Y = g["Y"]
        mask = g["mask_r_1d"]
        yh, yl, se = split_hankel_fp16(np.ascontiguousarray(Y.T))
        self.YTh, self.YTl, self.yt_se = torch.from_numpy(yh).to(dev), torch.from_numpy(yl).to(dev), se
        yh, yl, se = split_hankel_fp16(np.ascontiguousarray(Y.T * mask[None, :]))
        self.YMh, self.YMl, self.ym_se = torch.from_numpy(yh).to(dev), torch.from_numpy(yl).to(dev), se
        mt = self.mx.m
        self.phase = mt.phase_half(self.dz)
        self.cphase = torch.conj(self.phase).resolve_conj().contiguous()
        self.nlp = integ.nl.params(self.dz)
        shp = (self.Nr, self.Nt)
        c64 = dict(dtype=torch.complex64, device=dev)
        self.gam = torch.zeros(shp, **c64)
        self.a1 = torch.zeros(shp, **c64)
        self.g4 = torch.zeros(shp, **c64)
        self.grho = torch.zeros(shp, dtype=torch.float32, device=dev)
        self.ste = bool(integ.cfg.enable_ste)
        # dL/drho_s: written by the nonlinear adjoint when the forward applies the STE
        # index term, otherwise it stays zero (the scan's adjoint still reads it)
        self.grho_s = torch.zeros(shp, dtype=torch.float32, device=dev) if self.ste else None
        self.scratch = torch.zeros((4,) + shp, **c64)
        self.aslots = torch.zeros(4, dtype=torch.int32, device=dev)
        self.dt_invE2 = float(g["dt"] * g["invE2"])
        self._ckpts = None
        self._n = 0

    # ------------------------------------------------------------------ helpers
    def _aslot(self, i):
        return ctypes.c_void_p(self.aslots.data_ptr() + 4 * i)

    def _gemm(self, Yh, Yl, se, B, C, alpha, phase, bmax, cmax, reset):
        call("fp_hankel_gemm_mixed", ptr(Yh), ptr(Yl), se, ptr(B), ptr(C), self.Nr, self.Nt,
             ctypes.c_float(alpha), None, ptr(phase), bmax, cmax, reset, stream_ptr())

    def set_input(self, u0):
        """Load the time-domain input field (Nr x Nt, any complex dtype)."""
        self.mx.set_time(u0)

    def _step(self):
        it = self.it
        self.mx.step(self.dz, it.rho, it.rho_s, it._rho_s_arg(), self.nlp)

    def fluence(self):
        """F(r) of the current state (float64, Parseval over the spectral rows)."""
        s = self.mx.spec
        return (s.real.double() ** 2 + s.imag.double() ** 2).sum(dim=1) * (self.dt_invE2 / self.Nt)

    # ------------------------------------------------------------------ forward / backward
    def forward(self, n):
        """Run n steps from the loaded input, keeping checkpoints; returns F(r)."""
        K = self.ckpt or max(1, int(math.ceil(math.sqrt(n))))
        self._K, self._n = K, n
        self._ckpts = {}
        for i in range(n):
            if i % K == 0:
                self._ckpts[i] = self.mx.spec.clone()
            self._step()
        return self.fluence()

    def _forward_keep(self, finish):
        """One forward step from mx.spec, keeping what its adjoint needs: the state
        (spectral), its carrier densities (rho, and rho_s with STE) and v1 = the field
        after the first half linear step (spectral). With finish=False the step stops
        after v1."""
        mx, it, st = self.mx, self.it, stream_ptr()
        mt = mx.m
        u = mx.spec.clone()
        mx.plasma(it.rho, it.rho_s)                                      # clears slots 1-3
        self._gemm(mt.Yh, mt.Yl, mt.y_scale_exp, mx.spec, mx.w1, mt.Rk, self.phase,
                   mx._slot(0), mx._slot(1), None)
        self._gemm(mt.Yh, mt.Yl, mt.y_scale_exp, mx.w1, mx.w0, mt.iRk, None,
                   mx._slot(1), None, mx._slot(0))
        keep = (u, it.rho.clone(), mx.w0.clone(), it.rho_s.clone() if self.ste else None)
        if finish:
            call("fp_nl_step_c64", ptr(mx.w0), ptr(mx.w0), ptr(it.rho), ptr(it._rho_s_arg()), 1, 1,
                 ctypes.byref(self.nlp), ctypes.byref(mt.tables), mx._slot(2), st)
            self._gemm(mt.Yh, mt.Yl, mt.y_scale_exp, mx.w0, mx.w1, mt.Rk, self.phase,
                       mx._slot(2), mx._slot(3), None)
            call("fp_hankel_gemm_mixed", ptr(mt.Yh), ptr(mt.Yl), mt.y_scale_exp, ptr(mx.w1),
                 ptr(mx.spec), self.Nr, self.Nt, ctypes.c_float(mt.iRk), ptr(mt.mask_r), None,
                 mx._slot(3), mx._slot(0), None, st)
        return keep

    def _adjoint_step(self, un, rho_n, v1_n, rhos_n):
        """gam <- gradient w.r.t. the state un (spectral) given gam w.r.t. step(un);
        rho_n, v1_n, rhos_n: that step's carrier densities, field after the first half
        linear step (spectral) and STE density (None with STE off)."""
        it, st = self.it, stream_ptr()
        t = it.t
        mt = self.mx.m
        self.aslots[1:].zero_()
        self._gemm(self.YMh, self.YMl, self.ym_se, self.gam, self.a1, mt.iRk, self.cphase,
                   self._aslot(0), self._aslot(1), None)
        self._gemm(self.YTh, self.YTl, self.yt_se, self.a1, self.g4, mt.Rk, None,
                   self._aslot(1), None, self._aslot(0))
        rs_arg = rhos_n if it._rho_s_arg() is not None else None
        call("fp_adj_nl_step_c64", ptr(v1_n), ptr(rho_n), ptr(rs_arg), ptr(self.g4),
             ptr(self.g4), ptr(self.grho), ptr(self.grho_s if rs_arg is not None else None),
             ptr(self.scratch), ctypes.byref(self.nlp), ctypes.byref(mt.tables),
             self._aslot(2), st)
        self._gemm(self.YTh, self.YTl, self.yt_se, self.g4, self.a1, mt.iRk, self.cphase,
                   self._aslot(2), self._aslot(3), None)
        self._gemm(self.YTh, self.YTl, self.yt_se, self.a1, self.gam, mt.Rk, None,
                   self._aslot(3), None, None)
        call("fp_adj_plasma_c64", ptr(un), ptr(rho_n), ptr(rhos_n), ptr(self.grho),
             ptr(self.grho_s), ptr(self.gam), ctypes.byref(t.plasma_params), ptr(t.W_lut),
             ptr(t.Ws_lut if self.ste else None), ptr(mt.tw), self._aslot(0), st)

    def backward(self, dF):
        """Gradient w.r.t. the input field (time domain, complex64) of a loss whose
        gradient w.r.t. F(r) is dF, after forward(n)."""
        if self._ckpts is None:
            raise RuntimeError("backward() needs a preceding forward()")
        n, K = self._n, self._K
        dF = dF.to(device=self.gam.device, dtype=torch.float64).reshape(-1)
        c = (2.0 * self.dt_invE2 * dF).to(torch.float32)
        self.gam.copy_(c[:, None] * self.mx.spec)
        self.aslots.zero_()
        call("fp_absmax_f32", ptr(self.gam), 2 * self.Nr * self.Nt, self._aslot(0), stream_ptr())
        for s0 in sorted(self._ckpts, reverse=True):
            s1 = min(s0 + K, n)
            self.mx.spec.copy_(self._ckpts[s0])
            self.mx.slots.zero_()
            call("fp_absmax_f32", ptr(self.mx.spec), 2 * self.Nr * self.Nt, self.mx._slot(0),
                 stream_ptr())
            kept = [self._forward_keep(finish=(i + 1 < s1)) for i in range(s0, s1)]
            for i in range(s1 - 1, s0 - 1, -1):
                self._adjoint_step(*kept[i - s0])
            del kept
        self._ckpts = None
        out = torch.empty_like(self.gam)
        call("fp_fft_rows_c64", ptr(self.gam), ptr(out), self.Nr, self.Nt, 1, ptr(self.mx.m.tw),
             stream_ptr())
        return out


class FluenceFromInput(torch.autograd.Function):
    """F(r) after n z-steps as a differentiable function of the input field."""

    @staticmethod
    def forward(ctx, u0, engine, n):
        engine.set_input(u0.detach())
        F = engine.forward(int(n))
        ctx.engine = engine
        return F

    @staticmethod
    def backward(ctx, dF):
        g = ctx.engine.backward(dF)
        return g, None, None


def fluence_from_input(u0, engine, n):
    return FluenceFromInput.apply(u0, engine, n)


# ---------------------------------------------------------------------- LG input
def laguerre(p, x):
    """L_p(x) by the three-term recurrence."""
    l0 = torch.ones_like(x)
    if p == 0:
        return l0
    l1 = 1.0 - x
    for k in range(1, p):
        l0, l1 = l1, ((2 * k + 1 - x) * l1 - k * l0) / (k + 1)
    return l1


def lg_modes(r, w0, P):
    """G_p(r) = L_p(2 r^2 / w0^2) exp(-r^2 / w0^2), p < P (P x Nr, float64)."""
    r = torch.as_tensor(r, dtype=torch.float64)
    x = 2.0 * r ** 2 / w0 ** 2
    return torch.stack([laguerre(p, x) * torch.exp(-r ** 2 / w0 ** 2) for p in range(P)])


def lg_field(theta, G, E0, tenv):
    """u0 = E0 A(r) exp(i phi(r)) tenv(t), A = a @ G, phi = b @ G, theta = (a, b)."""
    P = G.shape[0]
    a, b = theta[:P], theta[P:]
    A = a @ G
    ph = b @ G
    radial = E0 * A * torch.exp(1j * ph)
    return (radial[:, None] * tenv[None, :]).to(torch.complex64)


def time_envelope(g, cfg, device):
    tp = cfg.delta_t / math.sqrt(2 * math.log(2))
    return torch.exp(-torch.as_tensor(g["tlist"], dtype=torch.float64, device=device) ** 2 / tp ** 2)


def radial_weights(g, device):
    r = np.asarray(g["rlist"])
    return torch.as_tensor(2.0 * np.pi * r * np.diff(r, prepend=0.0), device=device)


def fluence_loss(F, F_target, w):
    return torch.sum(w * (F - F_target) ** 2) / torch.sum(w * F_target ** 2)


def fit_lg(engine, F_target, theta0, n_steps, iters=20, P=None, lr=1.0, history=None):
    """Fit the LG coefficients so the output fluence matches F_target (L-BFGS on the
    weighted L2 mismatch, gradients from the adjoint)."""
    it = engine.it
    dev = it.device
    P = P or theta0.numel() // 2
    G = lg_modes(it.g["rlist"], it.cfg.w0, P).to(dev)
    tenv = time_envelope(it.g, it.cfg, dev)
    w = radial_weights(it.g, dev)
    E0 = float(it.cfg.E0)
    theta = theta0.detach().clone().to(dev, torch.float64).requires_grad_(True)
    opt = torch.optim.LBFGS([theta], lr=lr, max_iter=iters, history_size=10,
                            line_search_fn="strong_wolfe")
    hist = [] if history is None else history

    def closure():
        opt.zero_grad()
        F = fluence_from_input(lg_field(theta, G, E0, tenv), engine, n_steps)
        loss = fluence_loss(F, F_target, w)
        loss.backward()
        hist.append(float(loss.detach()))
        return loss
    opt.step(closure)
    return theta.detach(), hist


def forward_fluence(integ, u0, n):
    """F(r) after n z-steps from the time-domain input u0, on either precision path
    (no checkpoints, no gradient): the reference for finite differences."""
    integ.u = u0
    for _ in range(n):
        integ._pipeline_step(integ.dz)
    if integ._mx is not None:
        s = integ._mx.spec_rows
        dt_invE2 = float(integ.g["dt"] * integ.g["invE2"])
        return (s.real.double() ** 2 + s.imag.double() ** 2).sum(dim=1) * (dt_invE2 / integ.Nt)
    s = integ._spec
    return (s.real ** 2 + s.imag ** 2).sum(dim=1) * (float(integ.g["dt"] * integ.g["invE2"]) / integ.Nt)


def lg_problem(P=16, seed=0, amp_sigma=0.1, phase_sigma=0.3):
    """Start point (the plain Gaussian: a_0 = 1, others 0) and hidden coefficients of a
    fit: a_p = delta_p0 + amp_sigma xi / (1 + p), b_p = phase_sigma xi' / (1 + p),
    xi, xi' standard normal from numpy.random.default_rng(seed)."""
    rng = np.random.default_rng(seed)
    damp = 1.0 / (1.0 + np.arange(P))
    theta0 = np.zeros(2 * P)
    theta0[0] = 1.0
    hidden = theta0.copy()
    hidden[:P] += amp_sigma * rng.standard_normal(P) * damp
    hidden[P:] += phase_sigma * rng.standard_normal(P) * damp
    return theta0, hidden
