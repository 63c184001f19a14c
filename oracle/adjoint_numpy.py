"""FP64 discrete adjoint of the split-step propagator -- TEST INFRASTRUCTURE ONLY.

The reference (D3/sim) has no adjoint; this module differentiates the oracle's
own forward model (``reference_numpy``: ``Integrator.advance`` = plasma scan,
then ``Integrator.step``, D3/sim/integrator.py:138-162) exactly, step by step,
in the "real gradient" convention used throughout: for a real loss L of a complex
field u, g = dL/dRe(u) + i dL/dIm(u), so that dL = Re(sum conj(g) du). For a
complex-linear map v = A u the adjoint is g_u = A^H g_v.

Pieces (each is checked against central finite differences of the forward piece
in the tests, so a sign, index or factor slip in the derivation fails there):

* ``half_linear_adj``: adjoint of LinearOperator.half_linear (operators.py:47-53),
  Rk iRk Y^T ifft(conj(exp(i phi) mask) fft(Y^T g)).
* ``alpha_half_adj``: adjoint of u -> u exp(-alpha(|u|^2, rho) dz/2)
  (integrator.py:142,146).
* ``rhs_adj``: adjoint of NonlinearOperator.split's rhs (operators.py:89-146) with
  respect to u and rho (Kerr with Raman convolution, photo-ionisation loss with
  the PCHIP derivative, plasma absorption / defocusing, the alpha u term). The
  STE term is not differentiated (STE off).
* ``rk4_adj``: discrete adjoint of the RK4 stage sequence (integrator.py:143-145).
* ``plasma_adj``: adjoint of the rate-equation recurrence (kernels.py:36-106, STE
  off) with respect to the intensity, by the backward linear recurrence
  lambda_k = g_k + (d n_{k+1} / d n_k) lambda_{k+1}.
* ``Adjoint.gradient``: a whole run of n steps and its gradient with respect to
  the input field, for the fluence-mismatch loss of ``fluence_loss``.
"""
from __future__ import annotations

import numpy as np

from . import reference_numpy as O

LN10 = np.log(10.0)


def laguerre(p, x):
    """L_p(x) by the three-term recurrence (k+1) L_{k+1} = (2k+1-x) L_k - k L_{k-1}."""
    l0 = np.ones_like(x)
    if p == 0:
        return l0
    l1 = 1.0 - x
    for k in range(1, p):
        l0, l1 = l1, ((2 * k + 1 - x) * l1 - k * l0) / (k + 1)
    return l1


def lg_modes(r, w0, P):
    """Radial Laguerre-Gauss modes G_p(r) = L_p(2 r^2 / w0^2) exp(-r^2 / w0^2), p < P."""
    x = 2.0 * r ** 2 / w0 ** 2
    return np.stack([laguerre(p, x) * np.exp(-r ** 2 / w0 ** 2) for p in range(P)])


def lg_field(theta, g, cfg, P):
    """u0(r, t) = E0 A(r) exp(i phi(r)) exp(-t^2 / tp^2), A = sum a_p G_p, phi = sum b_p G_p,
    theta = (a_0..a_{P-1}, b_0..b_{P-1})."""
    G = lg_modes(g["rlist"], cfg.w0, P)
    a, b = theta[:P], theta[P:]
    A = a @ G
    ph = b @ G
    tp = cfg.delta_t / np.sqrt(2 * np.log(2))
    return (cfg.E0 * A * np.exp(1j * ph))[:, None] * np.exp(-g["tlist"] ** 2 / tp ** 2)[None, :]


def lg_field_grad(theta, g, cfg, P, g0):
    """dL/dtheta from the field gradient g0 (real-gradient convention)."""
    G = lg_modes(g["rlist"], cfg.w0, P)
    a, b = theta[:P], theta[P:]
    A = a @ G
    ph = b @ G
    tp = cfg.delta_t / np.sqrt(2 * np.log(2))
    tt = np.exp(-g["tlist"] ** 2 / tp ** 2)
    s = np.sum(np.conj(g0) * tt[None, :], axis=1)            # per r
    base = cfg.E0 * np.exp(1j * ph)
    ga = np.real(s * base) @ G.T
    gb = np.real(s * 1j * base * A) @ G.T
    return np.concatenate([ga, gb])


def fluence(u, g):
    """Output-plane fluence F(r) = sum_t |u|^2 dt invE2 (J/cm^2 up to the 1e4 of _record)."""
    return np.sum(np.abs(u) ** 2, axis=1) * g["dt"] * g["invE2"]


def radial_weights(g):
    r = g["rlist"]
    return 2.0 * np.pi * r * np.diff(r, prepend=0.0)


def fluence_loss(u, g, F_target):
    """L = sum_r w_r (F - F*)^2 / sum_r w_r F*^2 and its field gradient."""
    w = radial_weights(g)
    F = fluence(u, g)
    Z = np.sum(w * F_target ** 2)
    L = float(np.sum(w * (F - F_target) ** 2) / Z)
    c = 4.0 * w * (F - F_target) / Z * g["dt"] * g["invE2"]
    return L, c[:, None] * u


class Adjoint:
    """Forward pieces and their adjoints on an oracle Integrator's operators."""

    def __init__(self, integ: O.Integrator):
        self.it = integ
        self.cfg, self.g = integ.cfg, integ.g
        self.lin, self.nl = integ.lin, integ.nl
        self.dz, self.dt = integ.dz, integ.dt
        nl = self.nl
        if nl.enable_ste and nl.ste_pref:
            raise NotImplementedError("the adjoint covers STE off")
        self.mA = nl.T_op ** 2 * nl.inv_U_nl
        self.mB = -nl.T_op * nl.inv_U_nl
        self.mC = nl.inv_U_nl
        self.C_ph = 1e6 * nl.Ui / (nl.n0 * O.c * O.epsilon_0)
        pc = ((1.0 if nl.en_plasma_absorb else 0.0)
              + ((nl.plasma_phase - 1.0) if nl.en_plasma_defoc else 0.0))
        self.q = -nl.plasma_pref * pc                        # extra = q rho u
        self.mask_r = integ.mask_r

    # ------------------------------------------------------------------ forward pieces
    def plasma(self, u):
        rho = np.zeros(u.shape)
        rho_s = np.zeros(u.shape)
        self.nl.update_plasma(u, rho, rho_s, self.dt)
        return rho

    def alpha_half(self, v, rho):
        return v * np.exp(-0.5 * self.dz * self.nl.alpha_only(v, rho))

    def rk4(self, v, rho):
        dz, nl = self.dz, self.nl
        ys, ks = [], []
        y = v
        for cst in (0.5, 0.5, 1.0, None):
            ys.append(y)
            k, _ = nl.split(y, rho)
            ks.append(k)
            if cst is not None:
                y = v + cst * dz * k
        out = v + dz / 6 * (ks[0] + 2 * ks[1] + 2 * ks[2] + ks[3])
        return out, ys

    def step(self, u):
        """One propagate() iteration without recording: plasma scan + Strang step."""
        rho = self.plasma(u)
        v1 = self.lin.half_linear(u, self.dz)
        v2 = self.alpha_half(v1, rho)
        v3, _ = self.rk4(v2, rho)
        v4 = self.alpha_half(v3, rho)
        return self.lin.half_linear(v4, self.dz) * self.mask_r

    def run(self, u0, n):
        u = u0.astype(np.complex128)
        for _ in range(n):
            u = self.step(u)
        return u

    # ------------------------------------------------------------------ derivatives
    def photo_derivs(self, a2, rho):
        """photo (operators.py:92-102), d photo / d|u|^2 and d photo / d rho."""
        nl = self.nl
        z = np.zeros_like(a2)
        if not nl.en_pi_loss:
            return z, z, z
        x = np.clip(a2 * nl.invE2, nl.Imin, nl.Imax)
        f = nl.f_spline(x)
        fp = nl.f_spline(x, 1)
        W = np.nan_to_num(np.abs(f), nan=0.0, posinf=0.0, neginf=0.0)
        dW = np.where((a2 * nl.invE2 > nl.Imin) & (a2 * nl.invE2 < nl.Imax), np.sign(f) * fp, 0.0)
        dW = np.nan_to_num(dW, nan=0.0, posinf=0.0, neginf=0.0)
        dl = 1.0 - rho / nl.rho_max
        depl = np.clip(dl, 0.0, 1.0)
        ddepl = np.where((dl > 0.0) & (dl < 1.0), -1.0 / nl.rho_max, 0.0)
        den = a2 + 1e-30
        photo = W * self.C_ph * depl / den
        d_a2 = self.C_ph * depl * (dW * nl.invE2 / den - W / den ** 2)
        d_rho = W * self.C_ph * ddepl / den
        return photo, d_a2, d_rho

    def alpha_derivs(self, a2, rho):
        nl = self.nl
        photo, pa, pr = self.photo_derivs(a2, rho)
        alpha = photo + (nl.plasma_pref * rho if nl.en_plasma_absorb else 0.0)
        alpha = np.nan_to_num(alpha, nan=0.0, posinf=0.0, neginf=0.0)
        on = alpha > 0.0
        da2 = np.where(on, pa, 0.0)
        drho = np.where(on, pr + (nl.plasma_pref if nl.en_plasma_absorb else 0.0), 0.0)
        return np.maximum(alpha, 0.0), da2, drho

    def alpha_half_adj(self, v, rho, gw):
        """Adjoint of w = v exp(-alpha dz/2): returns (g_v, g_rho)."""
        a2 = np.abs(v) ** 2
        alpha, da2, drho = self.alpha_derivs(a2, rho)
        e = np.exp(-0.5 * self.dz * alpha)
        s = np.real(np.conj(gw) * v) * e * (-0.5 * self.dz)
        return e * gw + 2.0 * s * da2 * v, s * drho

    def rhs_adj(self, y, rho, lam):
        """Adjoint of split()'s rhs at y: returns (g_y, g_rho) for output gradient lam."""
        nl = self.nl
        a2 = np.abs(y) ** 2
        gp = np.where(a2 < 1e-30, 0.0, lam)
        G = np.fft.fft(gp, axis=1)
        ga = np.fft.ifft(self.mA * G, axis=1)
        gb = np.fft.ifft(self.mB * G, axis=1)
        gc = np.fft.ifft(self.mC * G, axis=1)
        photo, pa, pr = self.photo_derivs(a2, rho)
        alpha, aa, ar = self.alpha_derivs(a2, rho)
        K = np.zeros_like(a2)
        if nl.en_kerr_inst:
            K = K + (1.0 - nl.f_R) * a2
        if nl.en_kerr_raman:
            K = K + nl.f_R * np.fft.ifft(np.fft.fft(a2, axis=1) * nl.R_f, axis=1).real
        kap = nl.kerr_pref
        gy = -1j * kap * K * ga + photo * gb + np.conj(self.q) * rho * gc + alpha * gp
        gK = -kap * np.imag(np.conj(ga) * y)
        gI = np.zeros_like(a2)
        if nl.en_kerr_inst:
            gI = gI + (1.0 - nl.f_R) * gK
        if nl.en_kerr_raman:
            gI = gI + nl.f_R * np.fft.ifft(np.conj(nl.R_f) * np.fft.fft(gK, axis=1), axis=1).real
        gI = gI + pa * np.real(np.conj(gb) * y) + aa * np.real(np.conj(gp) * y)
        grho = (pr * np.real(np.conj(gb) * y) + np.real(np.conj(gc) * self.q * y)
                + ar * np.real(np.conj(gp) * y))
        return gy + 2.0 * gI * y, grho

    def rk4_adj(self, ys, rho, gout):
        """Discrete adjoint of out = v + dz/6 (k1 + 2 k2 + 2 k3 + k4), k_s = rhs(y_s),
        y1 = v, y2 = v + dz/2 k1, y3 = v + dz/2 k2, y4 = v + dz k3."""
        dz = self.dz
        lk4 = dz / 6 * gout
        m4, r4 = self.rhs_adj(ys[3], rho, lk4)
        lk3 = dz / 3 * gout + dz * m4
        m3, r3 = self.rhs_adj(ys[2], rho, lk3)
        lk2 = dz / 3 * gout + 0.5 * dz * m3
        m2, r2 = self.rhs_adj(ys[1], rho, lk2)
        lk1 = dz / 6 * gout + 0.5 * dz * m2
        m1, r1 = self.rhs_adj(ys[0], rho, lk1)
        return gout + m1 + m2 + m3 + m4, r1 + r2 + r3 + r4

    def half_linear_adj(self, gv):
        lin = self.lin
        phase = (lin.delta_k - lin.rhorho ** 2 / (2 * lin.komega) * lin.inv_U_op) * self.dz / 2
        D = np.exp(1j * phase) * lin.spec_mask
        Yt = lin.Y.T.astype(np.complex128)
        x = np.fft.fft(lin.iRk * np.dot(Yt, gv), axis=1)
        return lin.Rk * np.dot(Yt, np.fft.ifft(np.conj(D) * x, axis=1))

    def plasma_adj(self, u, rho, grho):
        """Adjoint of rho = update_plasma(u) (STE off) with respect to u."""
        nl, dt = self.nl, self.dt
        kel = nl._kel
        size, lmin, istep = kel["NLUT"], kel["logImin"], kel["inv_dlog"]
        Imin, Imax, T = kel["Imin"], kel["Imax"], kel["W_LUT"]
        na, beta, itr = nl.rho_max, nl.avalanche, nl.inv_taur
        I = (u.real ** 2 + u.imag ** 2) * nl.invE2
        W = O.lut_rate(I, T, size, lmin, istep, Imin, Imax)
        Ic = np.where(~np.isfinite(I) | (I < 0.0), 0.0, I)
        Icl = np.minimum(np.maximum(Ic, Imin), Imax)
        fidx = (np.log10(Icl) - lmin) * istep
        idx = np.clip(fidx.astype(np.int64), 0, size - 2)
        r = T[idx] + (fidx - idx) * (T[idx + 1] - T[idx])
        dW = np.where((Ic > Imin) & (Ic < Imax) & np.isfinite(r) & (r >= 0.0),
                      (T[idx + 1] - T[idx]) * istep / (Icl * LN10), 0.0)
        Nr, Nt = u.shape
        lam = grho[:, Nt - 1].copy()
        gI = np.zeros_like(I)
        for k in range(Nt - 2, -1, -1):
            n = rho[:, k]
            Ia = 0.5 * (I[:, k] + I[:, k + 1])
            Wa = 0.5 * (W[:, k] + W[:, k + 1])
            dl = 1.0 - n / na
            depl = np.clip(dl, 0.0, 1.0)
            dd = np.where((dl > 0.0) & (dl < 1.0), -1.0 / na, 0.0)
            S = Wa * depl
            L = beta * Ia * depl - itr
            z = L * dt
            with np.errstate(over="ignore", invalid="ignore", divide="ignore"):
                e = np.exp(z)
                big = np.abs(z) > 1e-6
                zs = np.where(big, z, 1.0)
                phi1 = np.where(big, (e - 1.0) / zs, 1.0 + z * (0.5 + z * (1.0 / 6.0 + z / 24.0)))
                dphi1 = np.where(big, (z * e - e + 1.0) / zs ** 2, 0.5 + z / 3.0 + z * z / 8.0)
                xn = e * n + S * dt * phi1
            ok = np.isfinite(xn) & (xn >= 0.0) & (xn <= na)
            dxdL = dt * e * n + S * dt * dt * dphi1
            dxdS = dt * phi1
            a = np.where(ok, e + dxdS * Wa * dd + dxdL * beta * Ia * dd, 0.0)
            p = np.where(ok, dxdL * beta * depl, 0.0)     # d n_{k+1} / d Ia
            qq = np.where(ok, dxdS * depl, 0.0)           # d n_{k+1} / d Wa
            gI[:, k] += 0.5 * lam * (p + qq * dW[:, k])
            gI[:, k + 1] += 0.5 * lam * (p + qq * dW[:, k + 1])
            lam = grho[:, k] + a * lam
        return 2.0 * nl.invE2 * gI * u

    def step_adj(self, u, gnext):
        """Gradient with respect to u_n of a loss depending on u_{n+1} = step(u_n)."""
        rho = self.plasma(u)
        v1 = self.lin.half_linear(u, self.dz)
        v2 = self.alpha_half(v1, rho)
        v3, ys = self.rk4(v2, rho)
        g5 = gnext * self.mask_r
        g4 = self.half_linear_adj(g5)
        g3, gr = self.alpha_half_adj(v3, rho, g4)
        g2, gr2 = self.rk4_adj(ys, rho, g3)
        g1, gr3 = self.alpha_half_adj(v1, rho, g2)
        gu = self.half_linear_adj(g1)
        return gu + self.plasma_adj(u, rho, gr + gr2 + gr3)

    def gradient(self, u0, n, F_target):
        """Loss of the output fluence after n steps and its gradient w.r.t. u0
        (states of all steps kept: small cases only)."""
        us = [u0.astype(np.complex128)]
        for _ in range(n):
            us.append(self.step(us[-1]))
        L, gN = fluence_loss(us[-1], self.g, F_target)
        gcur = gN
        for k in range(n - 1, -1, -1):
            gcur = self.step_adj(us[k], gcur)
        return L, gcur
