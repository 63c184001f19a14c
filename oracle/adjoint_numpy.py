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
  respect to u, rho and rho_s (Kerr with Raman convolution, photo-ionisation loss
  with the PCHIP derivative, plasma absorption / defocusing, the STE index term
  i ste_pref rho_s u of operators.py:136-140, the alpha u term).
* ``rk4_adj``: discrete adjoint of the RK4 stage sequence (integrator.py:143-145).
* ``plasma_adj``: adjoint of the rate-equation recurrence (kernels.py:36-106) with
  respect to the intensity, by the backward linear recurrence
  lambda_k = g_k + J_k^T lambda_{k+1} over the state (n_e, n_s): J_k is the 2 x 2
  Jacobian of one step (the coupled exact-exponential updates, their clamps at 0
  and the n_e + n_s <= n_a rescaling); without STE only its n_e entry is nonzero.
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
        """Carrier densities (rho, rho_s) of the field u (update_plasma; rho_s = 0
        without STE)."""
        rho = np.zeros(u.shape)
        rho_s = np.zeros(u.shape)
        self.nl.update_plasma(u, rho, rho_s, self.dt)
        return rho, rho_s

    def alpha_half(self, v, rho):
        return v * np.exp(-0.5 * self.dz * self.nl.alpha_only(v, rho))

    def rk4(self, v, rho, rho_s):
        dz, nl = self.dz, self.nl
        ys, ks = [], []
        y = v
        for cst in (0.5, 0.5, 1.0, None):
            ys.append(y)
            k, _ = nl.split(y, rho, rho_s)
            ks.append(k)
            if cst is not None:
                y = v + cst * dz * k
        out = v + dz / 6 * (ks[0] + 2 * ks[1] + 2 * ks[2] + ks[3])
        return out, ys

    def step(self, u):
        """One propagate() iteration without recording: plasma scan + Strang step."""
        rho, rho_s = self.plasma(u)
        v1 = self.lin.half_linear(u, self.dz)
        v2 = self.alpha_half(v1, rho)
        v3, _ = self.rk4(v2, rho, rho_s)
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

    def rhs_adj(self, y, rho, rho_s, lam):
        """Adjoint of split()'s rhs at y: returns (g_y, g_rho, g_rho_s) for output
        gradient lam."""
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
        sp = nl.ste_pref                                     # extra += i sp rho_s u
        cx = self.q * rho + 1j * sp * rho_s
        gy = -1j * kap * K * ga + photo * gb + np.conj(cx) * gc + alpha * gp
        gK = -kap * np.imag(np.conj(ga) * y)
        gI = np.zeros_like(a2)
        if nl.en_kerr_inst:
            gI = gI + (1.0 - nl.f_R) * gK
        if nl.en_kerr_raman:
            gI = gI + nl.f_R * np.fft.ifft(np.conj(nl.R_f) * np.fft.fft(gK, axis=1), axis=1).real
        gI = gI + pa * np.real(np.conj(gb) * y) + aa * np.real(np.conj(gp) * y)
        grho = (pr * np.real(np.conj(gb) * y) + np.real(np.conj(gc) * self.q * y)
                + ar * np.real(np.conj(gp) * y))
        grho_s = np.real(np.conj(gc) * (1j * sp) * y)
        return gy + 2.0 * gI * y, grho, grho_s

    def rk4_adj(self, ys, rho, rho_s, gout):
        """Discrete adjoint of out = v + dz/6 (k1 + 2 k2 + 2 k3 + k4), k_s = rhs(y_s),
        y1 = v, y2 = v + dz/2 k1, y3 = v + dz/2 k2, y4 = v + dz k3. Returns
        (g_v, g_rho, g_rho_s)."""
        dz = self.dz
        lk4 = dz / 6 * gout
        m4, r4, s4 = self.rhs_adj(ys[3], rho, rho_s, lk4)
        lk3 = dz / 3 * gout + dz * m4
        m3, r3, s3 = self.rhs_adj(ys[2], rho, rho_s, lk3)
        lk2 = dz / 3 * gout + 0.5 * dz * m3
        m2, r2, s2 = self.rhs_adj(ys[1], rho, rho_s, lk2)
        lk1 = dz / 6 * gout + 0.5 * dz * m2
        m1, r1, s1 = self.rhs_adj(ys[0], rho, rho_s, lk1)
        return gout + m1 + m2 + m3 + m4, r1 + r2 + r3 + r4, s1 + s2 + s3 + s4

    def half_linear_adj(self, gv):
        lin = self.lin
        phase = (lin.delta_k - lin.rhorho ** 2 / (2 * lin.komega) * lin.inv_U_op) * self.dz / 2
        D = np.exp(1j * phase) * lin.spec_mask
        Yt = lin.Y.T.astype(np.complex128)
        x = np.fft.fft(lin.iRk * np.dot(Yt, gv), axis=1)
        return lin.Rk * np.dot(Yt, np.fft.ifft(np.conj(D) * x, axis=1))

    @staticmethod
    def _exp_parts(x, S, L, dt):
        """exact_exp_step (kernels.py:28-34) with its pieces: (x_new, e, phi1, dphi1/dz)
        at z = L dt."""
        z = L * dt
        with np.errstate(over="ignore", invalid="ignore", divide="ignore"):
            e = np.exp(z)
            big = np.abs(z) > 1e-6
            zs = np.where(big, z, 1.0)
            phi1 = np.where(big, (e - 1.0) / zs, 1.0 + z * (0.5 + z * (1.0 / 6.0 + z / 24.0)))
            dphi1 = np.where(big, (z * e - e + 1.0) / zs ** 2, 0.5 + z / 3.0 + z * z / 8.0)
            xn = e * x + S * dt * phi1
        return xn, e, phi1, dphi1

    def _lut_d(self, I, table):
        """lut_rate (kernels.py:14-26) and its derivative d rate / d I (0 on the
        clamped ranges and where the rate is clamped to 0)."""
        kel = self.nl._kel
        size, lmin, istep = kel["NLUT"], kel["logImin"], kel["inv_dlog"]
        Imin, Imax = kel["Imin"], kel["Imax"]
        W = O.lut_rate(I, table, size, lmin, istep, Imin, Imax)
        Ic = np.where(~np.isfinite(I) | (I < 0.0), 0.0, I)
        Icl = np.minimum(np.maximum(Ic, Imin), Imax)
        fidx = (np.log10(Icl) - lmin) * istep
        idx = np.clip(fidx.astype(np.int64), 0, size - 2)
        r = table[idx] + (fidx - idx) * (table[idx + 1] - table[idx])
        dW = np.where((Ic > Imin) & (Ic < Imax) & np.isfinite(r) & (r >= 0.0),
                      (table[idx + 1] - table[idx]) * istep / (Icl * LN10), 0.0)
        return W, dW

    def recurrence_adj(self, I, W, Ws, rho, rho_s, grho, grho_s, rescale_jacobian=True):
        """Adjoint of rate_recurrence_rows (kernels.py:59-104) on given per-sample
        intensities I and rates W, Ws: gradients (g_I, g_W, g_Ws) for the density
        gradients grho, grho_s, where (rho, rho_s) are the recurrence's outputs.
        Step k maps (n_e, n_s)_k to (n_e, n_s)_{k+1}: a = E(n_e; S_e, L_e),
        b = E(n_s; S_s, L_s) (each clamped at 0), then (a, b) n_a / (a + b) where
        a + b > n_a. The backward sweep carries lambda = dL/d(n_e, n_s)_{k+1}, forms
        mu = dL/d(a, b) through the rescaling and the clamps, gives half of
        mu . d(a, b)/d(I_avg, W_avg, Ws_avg) to samples k and k + 1 each, and sets
        lambda_k = g_k + mu . d(a, b)/d(n_e, n_s)_k. rescale_jacobian=False takes the
        rescaling as the identity (a planted omission: the tests show the check
        against finite differences detects it)."""
        nl, dt = self.nl, self.dt
        ste = bool(nl.enable_ste)
        na, bg, bs = nl.rho_max, nl.avalanche, nl.avalanche_s
        itr, its = nl.inv_taur, nl.inv_tau_ste
        Nr, Nt = I.shape
        le = grho[:, Nt - 1].copy()
        ls = grho_s[:, Nt - 1].copy()
        gI, gW, gWs = np.zeros_like(I), np.zeros_like(I), np.zeros_like(I)
        z = np.zeros(Nr)
        for k in range(Nt - 2, -1, -1):
            ne, ns = rho[:, k], rho_s[:, k]
            Ia = 0.5 * (I[:, k] + I[:, k + 1])
            Wa = 0.5 * (W[:, k] + W[:, k + 1])
            Wsa = 0.5 * (Ws[:, k] + Ws[:, k + 1])
            dl = 1.0 - ne / na
            depl = np.clip(dl, 0.0, 1.0)
            dd = np.where((dl > 0.0) & (dl < 1.0), -1.0 / na, 0.0)
            wst = (Wsa + bs * Ia * ne) if ste else z
            Se = Wa * depl + (wst * ns / na if ste else 0.0)
            Le = bg * Ia * depl - itr
            a, ee, ph, dph = self._exp_parts(ne, Se, Le, dt)
            oka = np.isfinite(a) & (a >= 0.0)
            a = np.where(oka, a, 0.0)
            aS, aL = dt * ph, dt * ee * ne + Se * dt * dt * dph      # da/dS_e, da/dL_e
            a_ne = ee + aS * (Wa * dd + (bs * Ia * ns / na if ste else 0.0)) + aL * bg * Ia * dd
            a_ns = aS * wst / na
            a_I = aS * (bs * ne * ns / na if ste else 0.0) + aL * bg * depl
            a_W = aS * depl
            a_Ws = aS * ns / na
            if ste:
                Ss = itr * ne
                Ls = -wst / na - its
                b, es, phs, dphs = self._exp_parts(ns, Ss, Ls, dt)
                okb = np.isfinite(b) & (b >= 0.0)
                b = np.where(okb, b, 0.0)
                bS, bL = dt * phs, dt * es * ns + Ss * dt * dt * dphs
                b_ne = bS * itr - bL * bs * Ia / na
                b_ns = es
                b_I = -bL * bs * ne / na
                b_Ws = -bL / na
            else:
                b = z
                okb = np.zeros(Nr, bool)
                b_ne = b_ns = b_I = b_Ws = z
            tot = a + b
            over = (tot > na) & rescale_jacobian
            T2 = np.where(over, tot * tot, 1.0)
            mu_a = np.where(over, na * b / T2 * (le - ls), le)
            mu_b = np.where(over, na * a / T2 * (ls - le), ls)
            mu_a = np.where(oka, mu_a, 0.0)
            mu_b = np.where(okb, mu_b, 0.0)
            gIa = 0.5 * (mu_a * a_I + mu_b * b_I)
            gWa = 0.5 * mu_a * a_W
            gWsa = 0.5 * (mu_a * a_Ws + mu_b * b_Ws)
            gI[:, k] += gIa
            gI[:, k + 1] += gIa
            gW[:, k] += gWa
            gW[:, k + 1] += gWa
            gWs[:, k] += gWsa
            gWs[:, k + 1] += gWsa
            le, ls = (grho[:, k] + mu_a * a_ne + mu_b * b_ne,
                      grho_s[:, k] + mu_a * a_ns + mu_b * b_ns)
        return gI, gW, gWs

    def plasma_adj(self, u, rho, rho_s, grho, grho_s):
        """Adjoint of (rho, rho_s) = update_plasma(u) with respect to u, for the
        density gradients grho, grho_s: the recurrence's adjoint, the rates through
        their LUT derivatives, then dL/du = 2 invE2 (dL/dI) u."""
        nl = self.nl
        kel = nl._kel
        I = (u.real ** 2 + u.imag ** 2) * nl.invE2
        W, dW = self._lut_d(I, kel["W_LUT"])
        if nl.enable_ste:
            Ws, dWs = self._lut_d(I, kel["W_LUT_s"])
        else:
            Ws, dWs = np.zeros_like(I), np.zeros_like(I)
        gI, gW, gWs = self.recurrence_adj(I, W, Ws, rho, rho_s, grho, grho_s)
        return 2.0 * nl.invE2 * (gI + gW * dW + gWs * dWs) * u

    def step_adj(self, u, gnext):
        """Gradient with respect to u_n of a loss depending on u_{n+1} = step(u_n)."""
        rho, rho_s = self.plasma(u)
        v1 = self.lin.half_linear(u, self.dz)
        v2 = self.alpha_half(v1, rho)
        v3, ys = self.rk4(v2, rho, rho_s)
        g5 = gnext * self.mask_r
        g4 = self.half_linear_adj(g5)
        g3, gr = self.alpha_half_adj(v3, rho, g4)
        g2, gr2, grs = self.rk4_adj(ys, rho, rho_s, g3)
        g1, gr3 = self.alpha_half_adj(v1, rho, g2)
        gu = self.half_linear_adj(g1)
        return gu + self.plasma_adj(u, rho, rho_s, gr + gr2 + gr3, grs)

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
